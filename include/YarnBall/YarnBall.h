// YarnBall.h — drop-in C++ surface of the MI355X stepper.
//
// Source-compatible with the reference's public header
// (KittenEngine/YarnBall/YarnBall.h:16-206) for everything the simulation
// path uses: VertexFlags, Vertex (same 32-byte layout), MetaData (same field
// names), and class Sim with the same public fields and methods.  glm is not
// required: the small vec3/vec4/Rotor/Bound types below carry the same member
// names (x, y, z, w; min, max) and memory layout as the glm/Kitten types they
// stand in for.
//
// Every method forwards to the C ABI in yarnball_hip.h (libyarnball_hip.so).
// Errors are reported the way the reference reports them: std::runtime_error
// with the reference's messages, lastErrorCode / lastWarningCode, and a
// stderr warning when printErrors is set.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../yarnball_hip.h"

namespace YarnBall {

struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
    vec3 operator+(const vec3& o) const { return vec3(x + o.x, y + o.y, z + o.z); }
    vec3 operator-(const vec3& o) const { return vec3(x - o.x, y - o.y, z - o.z); }
    vec3 operator*(float s) const { return vec3(x * s, y * s, z * s); }
    vec3& operator+=(const vec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    vec3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
};
inline vec3 operator*(float s, const vec3& v) { return v * s; }

struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    vec4() = default;
    vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
};

}  // namespace YarnBall

namespace Kitten {
// Quaternion ("rotor") with bivector part {yz, zx, xy} in x, y, z and scalar w
// (KittenEngine/includes/modules/Rotor.h:7-20).
struct Rotor {
    float x = 0, y = 0, z = 0, w = 1;
    Rotor() = default;
    Rotor(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    static Rotor identity() { return Rotor(); }
};
// Axis-aligned box (KittenEngine/includes/modules/Bound.h:7-20).
template <int dim = 3, typename Real = float>
struct Bound {
    YarnBall::vec3 min = YarnBall::vec3(std::numeric_limits<float>::infinity());
    YarnBall::vec3 max = YarnBall::vec3(-std::numeric_limits<float>::infinity());
    YarnBall::vec3 center() const { return (min + max) * 0.5f; }
};
namespace LBVH {
using aabb = Bound<3, float>;
}
}  // namespace Kitten
namespace Kit = Kitten;

namespace YarnBall {

constexpr int MAX_COLLISIONS_PER_SEGMENT = YB_MAX_COLLISIONS_PER_SEGMENT;

enum class VertexFlags {
    hasPrev = YB_FLAG_HAS_PREV,
    hasNext = YB_FLAG_HAS_NEXT,
    hasNextOrientation = YB_FLAG_HAS_NEXT_ORIENTATION,
    fixOrientation = YB_FLAG_FIX_ORIENTATION,
    colliding = YB_FLAG_COLLIDING,
};

inline bool hasFlag(const uint32_t flags, const VertexFlags flag) {
    return (flags & (uint32_t)flag) != 0;
}
inline uint32_t setFlag(const uint32_t flags, const VertexFlags flag, const bool state) {
    return state ? flags | (uint32_t)flag : flags & ~(uint32_t)flag;
}

// Simulation vertex, YarnBall.h:52-61.
typedef struct {
    vec3 pos;
    float invMass;
    float lRest;
    float kStretch;
    int connectionIndex;
    uint32_t flags;
} Vertex;
static_assert(sizeof(Vertex) == sizeof(yb_vertex), "Vertex must stay 32 bytes");

// Simulation parameters, the scalar fields of YarnBall.h:63-114 in the C ABI's
// order (device pointers live inside the library).
struct MetaData {
    vec3 gravity;
    int numItr;
    vec3 worldFloorNormal;
    float worldFloorPos;
    float h;
    float lastH;
    float time;
    int numVerts;
    float damping;
    float drag;
    float frictionCoeff;
    float kCollision;
    float kFriction;
    float detectionRadius;
    float scaledDetectionRadius;
    float radius;
    float accelerationRatio;
    float barrierThickness;
    float detectionScaler;
    float bvhRebuildPeriod;
    int detectionPeriod;
    float maxSegLen;
    float minSegLen;
    int useStepSizeLimit;
    int useVelocityRadius;
};
static_assert(sizeof(MetaData) == sizeof(yb_params), "MetaData mirrors yb_params");

// Column-major 4x4 matrix (glm::mat4 layout: m[column][row]), for the
// transform argument of the curve readers.
struct mat4 {
    vec4 m[4];
    explicit mat4(float d = 1.f) {
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) m[c][r] = c == r ? d : 0.f;
    }
    vec4& operator[](int c) { return m[c]; }
    const vec4& operator[](int c) const { return m[c]; }
};

class Sim {
public:
    enum {
        ERROR_NONE = 0,
        ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED,
        WARNING_SEGMENT_INTERPENETRATION,
        ERROR_TRAVERSAL_STACK_EXCEEDED  // no counterpart in the reference (its LBVH is external)
    };

    Vertex* verts;
    Kit::Rotor* qs;
    vec4* qRests;
    vec3* vels;

    MetaData meta;
    float maxH = 1e-3f;
    std::vector<float> initialInvMasses;
    Kit::Bound<> currentBounds;

    int lastErrorCode = ERROR_NONE;
    int lastWarningCode = ERROR_NONE;

    bool printErrors = true;
    bool renderShaded = false;

    explicit Sim(int numVerts);
    ~Sim();
    Sim(const Sim&) = delete;
    Sim& operator=(const Sim&) = delete;

    size_t stepCount() { return (size_t)yb_step_count(handle); }

    void configure(float density = 1e-3f);
    void setKBend(float k = 3e-9f);
    void setKStretch(float k = 1e-2f);

    void upload();
    void download();
    void zeroVelocities();

    void step(float dt);
    float advance(float dt);

    void printCollisionStats();
    Kit::LBVH::aabb bounds();

    // Scene files (io/fileExport.cpp, io/jsonBuilder.cpp:174-212).
    void exportToBCC(std::string path, bool exportAsPolyline = false);
    void exportToOBJ(std::string path);
    void glueEndpoints(float searchRadius);

    // A Sim holding a yb_scene's arrays (configured: also created on the
    // device and uploaded, as buildFromJSON returns it).
    static Sim* fromScene(yb_scene* scene, bool configured);

    // Device handle, for callers that need the C ABI directly.
    yb_sim* deviceHandle() { return handle; }

private:
    yb_sim* handle = nullptr;
    void raise(int code, const char* where);
};

// Scene readers (io/reader.cpp:50-295, io/jsonBuilder.cpp:8-170).
Sim* readFromOBJ(std::string path, float targetSegLen, mat4 transform,
                 bool breakUpClosedCurves = false, bool allowResample = true);
Sim* readFromBCC(std::string path, float targetSegLen, mat4 transform,
                 bool breakUpClosedCurves = false, bool allowResample = true);
Sim* readFromPoly(std::string path, float targetSegLen, mat4 transform,
                  bool breakUpClosedCurves = false, bool allowResample = true);
Sim* buildFromJSON(std::string path);

}  // namespace YarnBall
