// scene_io.cpp -- host-side scene construction and export for the stepper:
// the JSON scene description, the curve file readers with Catmull-Rom
// resampling, endpoint gluing, and the BCC / OBJ writers.
//
// Reference behaviour restated here:
//   buildFromJSON            KittenEngine/YarnBall/io/jsonBuilder.cpp:8-170
//   Sim::glueEndpoints       jsonBuilder.cpp:174-212
//   createFromCurves         KittenEngine/YarnBall/io/reader.cpp:13-48
//   readFromBCC / OBJ / Poly reader.cpp:50-295
//   resampleCMR (contract)   KittenEngine/YarnBall/io/resample.h:10-188
//   cmrSpline (the curve)    KittenEngine/includes/modules/Common.h:681-694
//   exportToBCC / OBJ        KittenEngine/YarnBall/io/fileExport.cpp:6-100
//   BCCHeader                KittenEngine/YarnBall/YarnBall.h:23-32
//
// Deliberate differences: the JSON reader accepts the comments and trailing
// commas the shipped configs contain (jsoncpp's lenient mode); the next
// point of a closed .poly curve at its start vertex is the partner of the
// first segment listed for it (the reference takes whatever its hash map
// yields first); glueEndpoints visits endpoints in index order and finds the
// nearest vertex through a uniform grid instead of an O(E N) scan (the
// reference's OpenMP loop makes the order of competing glues a race); the
// resampler solves for equal chords over the whole curve at once (see
// equalize_chords) where the reference marches chords from both ends,
// splices the two marches and blends the seams, so vertex positions differ
// from the reference's by a fraction of a segment while counts and spacing
// agree.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cctype>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/yarnball_hip.h"

namespace yb {
void set_message(const std::string& m);  // capi.hip: the message behind yb_last_message()
}

namespace {

// ---------------------------------------------------------------------------
// JSON: a small DOM and a lenient recursive-descent reader.
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const {
        if (kind != Obj) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool isNull() const { return kind == Null; }
    bool isArray() const { return kind == Arr; }
    double asNum() const {
        if (kind == Num) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        throw std::runtime_error("JSON: expected a number");
    }
    float asFloat() const { return (float)asNum(); }
    int asInt() const { return (int)asNum(); }
    bool asBool() const {
        if (kind == Bool) return b;
        if (kind == Num) return num != 0.0;
        throw std::runtime_error("JSON: expected a boolean");
    }
    std::string asString() const {
        if (kind != Str) throw std::runtime_error("JSON: expected a string");
        return str;
    }
};

// A missing key reads as null, like jsoncpp's operator[] on a const value.
const JVal& member(const JVal& v, const char* key) {
    static const JVal kNull;
    const JVal* p = v.get(key);
    return p ? *p : kNull;
}

class JsonReader {
public:
    explicit JsonReader(const std::string& text) : s_(text) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char* what) {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); k++) line += s_[k] == '\n';
        throw std::runtime_error(std::string("Error parsing the JSON file: ") + what + " at line " +
                                 std::to_string(line));
    }
    void ws() {
        for (;;) {
            while (i_ < s_.size() && isspace((unsigned char)s_[i_])) i_++;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') i_++;
            } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 2;
            } else {
                return;
            }
        }
    }
    bool lit(const char* w) {
        const size_t l = strlen(w);
        if (s_.compare(i_, l, w) == 0) {
            i_ += l;
            return true;
        }
        return false;
    }
    JVal value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        const char c = s_[i_];
        JVal v;
        if (c == '{') {
            v.kind = JVal::Obj;
            i_++;
            for (;;) {
                ws();
                if (i_ < s_.size() && s_[i_] == '}') {
                    i_++;
                    break;
                }
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a key");
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                i_++;
                JVal m = value();
                v.obj.emplace_back(std::move(k), std::move(m));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                ws();
                if (i_ < s_.size() && s_[i_] == '}') {
                    i_++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::Arr;
            i_++;
            for (;;) {
                ws();
                if (i_ < s_.size() && s_[i_] == ']') {
                    i_++;
                    break;
                }
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                ws();
                if (i_ < s_.size() && s_[i_] == ']') {
                    i_++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JVal::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JVal::Bool;
        } else if (lit("null")) {
        } else {
            const char* b = s_.c_str() + i_;
            char* e = nullptr;
            v.num = strtod(b, &e);
            if (e == b) fail("unexpected character");
            v.kind = JVal::Num;
            i_ += (size_t)(e - b);
        }
        return v;
    }
    std::string string() {
        i_++;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                const char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        const unsigned cp = (unsigned)strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;  // \" \\ \/
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        i_++;
        return out;
    }
};

// ---------------------------------------------------------------------------
// Geometry helpers.
struct f3 {
    float x, y, z;
};
struct d3 {
    double x, y, z;
};
inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dlen2(d3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double dlen(d3 a) { return sqrt(dlen2(a)); }
inline d3 tod(f3 p) { return {(double)p.x, (double)p.y, (double)p.z}; }

// 3x4 affine map, row-major: p'_i = sum_j m[i][j] p_j + m[i][3], evaluated
// in float column by column like a 4x4 matrix times (p, 1).
struct Affine {
    float m[3][4] = {{0.01f, 0, 0, 0}, {0, 0.01f, 0, 0}, {0, 0, 0.01f, 0}};
    f3 apply(float x, float y, float z) const {
        f3 r;
        float* o = &r.x;
        for (int i = 0; i < 3; i++) o[i] = ((m[i][0] * x + m[i][1] * y) + m[i][2] * z) + m[i][3];
        return r;
    }
};

// ---------------------------------------------------------------------------
// Uniform Catmull-Rom spline through control points cp[0 .. n-1].  The curve
// runs from cp[1] (u = 1) to cp[n-2] (u = n-2); span k, k <= u < k+1,
// interpolates cp[k] -> cp[k+1] with end tangents (cp[k+1] - cp[k-1]) / 2
// and (cp[k+2] - cp[k]) / 2 (the curve and parameter range the reference
// samples, resample.h:10-18 over Common.h:681-694).
class CmrCurve {
public:
    explicit CmrCurve(const std::vector<f3>& cp) : cp_(cp), n_((int)cp.size()) {}
    double first() const { return 1.0; }
    double last() const { return n_ - 2.0; }
    int spans() const { return n_ - 3; }

    d3 at(double u) const {
        const int k = std::min(std::max((int)std::floor(u), 1), n_ - 3);
        const double s = u - k, s2 = s * s, s3 = s2 * s;
        const d3 p0 = tod(cp_[k - 1]), p1 = tod(cp_[k]), p2 = tod(cp_[k + 1]), p3 = tod(cp_[k + 2]);
        // cubic Hermite basis with the Catmull-Rom tangents
        return p1 * (2 * s3 - 3 * s2 + 1) + (p2 - p0) * (0.5 * (s3 - 2 * s2 + s)) +
               p2 * (3 * s2 - 2 * s3) + (p3 - p1) * (0.5 * (s3 - s2));
    }

    // dx/du, the derivative of the Hermite form above
    d3 tangent(double u) const {
        const int k = std::min(std::max((int)std::floor(u), 1), n_ - 3);
        const double s = u - k, s2 = s * s;
        const d3 p0 = tod(cp_[k - 1]), p1 = tod(cp_[k]), p2 = tod(cp_[k + 1]), p3 = tod(cp_[k + 2]);
        return (p2 - p1) * (6 * (s - s2)) + (p2 - p0) * (0.5 * (3 * s2 - 4 * s + 1)) +
               (p3 - p1) * (0.5 * (3 * s2 - 2 * s));
    }

    // Length of the Bezier control polygon of span k (points p1,
    // p1 + (p2 - p0) / 6, p2 - (p3 - p1) / 6, p2): an upper bound of the
    // span's arc length.
    double span_bound(int k) const {
        const d3 p0 = tod(cp_[k - 1]), p1 = tod(cp_[k]), p2 = tod(cp_[k + 1]), p3 = tod(cp_[k + 2]);
        const d3 b1 = p1 + (p2 - p0) * (1.0 / 6.0), b2 = p2 - (p3 - p1) * (1.0 / 6.0);
        return dlen(b1 - p1) + dlen(b2 - b1) + dlen(p2 - b2);
    }

private:
    const std::vector<f3>& cp_;
    int n_;
};

// Arc length of the curve: parameters tabulated at most segLen / 8 of arc
// apart (by the control-polygon bound), the arc between neighbours by
// 5-point Gauss-Legendre quadrature of |dx/du|, accumulated.
class ArcTable {
public:
    ArcTable(const CmrCurve& cv, double segLen) : cv_(cv) {
        for (int k = 1; k <= cv.spans(); k++) {
            const int m = (int)std::min(std::max(std::ceil(8.0 * cv.span_bound(k) / segLen), 1.0), 65536.0);
            for (int j = 0; j < m; j++) us_.push_back(k + j / (double)m);
        }
        us_.push_back(cv.last());
        s_.assign(us_.size(), 0.0);
        for (size_t i = 1; i < us_.size(); i++) s_[i] = s_[i - 1] + arc(us_[i - 1], us_[i]);
    }

    double length() const { return s_.back(); }

    // arc length from u0 to u1 inside one table interval
    double arc(double u0, double u1) const {
        static const double x[5] = {-0.9061798459386640, -0.5384693101056831, 0.0,
                                    0.5384693101056831, 0.9061798459386640};
        static const double w[5] = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889,
                                    0.4786286704993665, 0.2369268850561891};
        const double h = 0.5 * (u1 - u0), c = 0.5 * (u0 + u1);
        double sum = 0.0;
        for (int q = 0; q < 5; q++) sum += w[q] * dlen(cv_.tangent(c + h * x[q]));
        return sum * h;
    }

    // The parameter at arc length s from the start: the table interval
    // holding s, then bisection on the arc inside it.
    double param_at(double s) const {
        size_t i = (size_t)(std::upper_bound(s_.begin(), s_.end(), s) - s_.begin());
        i = std::min(std::max(i, (size_t)1), s_.size() - 1) - 1;
        double lo = us_[i], hi = us_[i + 1];
        const double rest = s - s_[i];
        for (int b = 0; b < 60; b++) {
            const double mid = 0.5 * (lo + hi);
            if (mid <= lo || mid >= hi) break;
            (arc(us_[i], mid) < rest ? lo : hi) = mid;
        }
        return 0.5 * (lo + hi);
    }

private:
    const CmrCurve& cv_;
    std::vector<double> us_, s_;
};

// Solves for parameters u_1 .. u_{m-1} (u_0 and u_m fixed) and a chord
// length c so that every chord |x(u_{k+1}) - x(u_k)| equals c: Gauss-Newton
// on the residuals r_k = |x(u_{k+1}) - x(u_k)| - c with Levenberg-Marquardt
// damping.  Each residual touches two neighbouring parameters and c, so the
// normal equations are tridiagonal with one dense border row and column,
// solved by two Thomas sweeps and a Schur complement for c.  Steps that
// would reorder the parameters, or do not lower the sum of squares, are
// retried with more damping.  Where the curve has a loop narrower than c no
// placement on it has equal chords; the solve then settles on the least-
// squares compromise (points stay on the loop, chords there a little short).
// Returns the largest |r_k| / c of the last iterate checked.
double equalize_chords(const CmrCurve& cv, std::vector<double>& u) {
    const int m = (int)u.size() - 1;
    std::vector<d3> x(m + 1), t(m + 1);
    std::vector<double> r(m), a(m), b(m);
    auto chords = [&](const std::vector<double>& uu, double& sum) {
        for (int k = 0; k <= m; k++) x[k] = cv.at(uu[k]);
        sum = 0.0;
        for (int k = 0; k < m; k++) sum += dlen(x[k + 1] - x[k]);
    };
    auto cost = [&](const std::vector<double>& uu, double cc) {
        double sq = 0.0;
        for (int k = 0; k < m; k++) {
            const double l = dlen(cv.at(uu[k + 1]) - cv.at(uu[k])) - cc;
            sq += l * l;
        }
        return sq;
    };
    double total = 0.0;
    chords(u, total);
    double c = total / m;
    if (m < 2 || c <= 0.0) return 0.0;
    const int nu = m - 1;  // unknown parameters u_1 .. u_{m-1}
    std::vector<double> diag(nu), off(nu), w(nu), g(nu), y(nu), z(nu), cp(nu), du(nu);
    std::vector<double> trial(m + 1);
    double lambda = 1e-4, worst = INFINITY;
    for (int it = 0; it < 100; it++) {
        double sq = 0.0;
        worst = 0.0;
        for (int k = 0; k <= m; k++) {
            x[k] = cv.at(u[k]);
            t[k] = cv.tangent(u[k]);
        }
        for (int k = 0; k < m; k++) {
            const d3 d = x[k + 1] - x[k];
            const double l = dlen(d);
            const d3 e = l > 0.0 ? d * (1.0 / l) : d3{0, 0, 0};
            r[k] = l - c;
            a[k] = -(e.x * t[k].x + e.y * t[k].y + e.z * t[k].z);       // d r_k / d u_k
            b[k] = e.x * t[k + 1].x + e.y * t[k + 1].y + e.z * t[k + 1].z;  // d r_k / d u_{k+1}
            sq += r[k] * r[k];
            worst = std::max(worst, fabs(r[k]));
        }
        worst /= c;
        if (worst <= 1e-10) break;
        // normal equations: unknown j (0-based) is u_{j+1}, touched by rows j+1 (as u_k) and j (as u_{k+1})
        double gc = 0.0, sc = (double)m;
        for (int k = 0; k < m; k++) gc -= r[k];
        for (int j = 0; j < nu; j++) {
            const double ak = a[j + 1], bk = b[j];
            diag[j] = ak * ak + bk * bk;
            off[j] = j + 1 < nu ? a[j + 1] * b[j + 1] : 0.0;  // (u_{j+1}, u_{j+2}) through row j+1
            w[j] = -ak - bk;
            g[j] = ak * r[j + 1] + bk * r[j];
        }
        bool accepted = false;
        for (int tries = 0; tries < 30 && !accepted; tries++) {
            // Thomas sweeps on T = tridiag(off, diag (1 + lambda), off) for -g and w
            double prevC = 0.0, prevY = 0.0, prevZ = 0.0;
            for (int j = 0; j < nu; j++) {
                const double dj = diag[j] * (1.0 + lambda) + 1e-300;
                const double lo = j > 0 ? off[j - 1] : 0.0;
                const double den = dj - lo * prevC;
                cp[j] = off[j] / den;
                y[j] = (-g[j] - lo * prevY) / den;
                z[j] = (w[j] - lo * prevZ) / den;
                prevC = cp[j];
                prevY = y[j];
                prevZ = z[j];
            }
            for (int j = nu - 2; j >= 0; j--) {
                y[j] -= cp[j] * y[j + 1];
                z[j] -= cp[j] * z[j + 1];
            }
            double wy = 0.0, wz = 0.0;
            for (int j = 0; j < nu; j++) {
                wy += w[j] * y[j];
                wz += w[j] * z[j];
            }
            const double dc = (-gc - wy) / (sc * (1.0 + lambda) - wz);
            bool ordered = true;
            trial[0] = u[0];
            trial[m] = u[m];
            for (int j = 0; j < nu; j++) {
                du[j] = y[j] - z[j] * dc;
                trial[j + 1] = u[j + 1] + du[j];
            }
            for (int k = 0; k < m && ordered; k++) ordered = trial[k + 1] > trial[k];
            const double sqNew = ordered ? cost(trial, c + dc) : INFINITY;
            if (sqNew < sq) {
                u.swap(trial);
                c += dc;
                lambda = std::max(lambda * 0.3, 1e-12);
                accepted = true;
                // a loop narrower than c leaves a least-squares remainder: stop once it settles
                if (sq - sqNew <= 1e-9 * sq) return worst;
            } else {
                lambda *= 10.0;
            }
        }
        if (!accepted) break;
    }
    return worst;
}

// Resamples the Catmull-Rom spline through cp to points about segLen apart,
// from cp[1] to cp[n-2] (resampleCMR, resample.h:149-188, has the same
// contract: consecutive points one straight length apart, both ends on the
// curve's ends).  The placement is this file's own: a first placement at
// equal arc length about segLen apart gives the sum S of its chords; the
// curve gets m = round(S / segLen) segments, starts from equal arc length
// L / m, and equalize_chords makes every chord the same length c ~ segLen.
std::vector<f3> resample_cmr(const std::vector<f3>& cp, double segLen) {
    const CmrCurve cv(cp);
    const ArcTable arc(cv, segLen);
    const double L = arc.length();
    auto equal_arc = [&](int m) {
        std::vector<double> u(m + 1);
        u[0] = cv.first();
        for (int k = 1; k < m; k++) u[k] = arc.param_at(L * k / m);
        u[m] = cv.last();
        return u;
    };
    const int m0 = std::max(1, (int)std::lround(L / segLen));
    const std::vector<double> u0 = equal_arc(m0);
    double chordSum = 0.0;
    for (int k = 0; k < m0; k++) chordSum += dlen(cv.at(u0[k + 1]) - cv.at(u0[k]));
    const int m = std::max(1, (int)std::lround(chordSum / segLen));
    std::vector<double> u = equal_arc(m);
    equalize_chords(cv, u);
    std::vector<f3> out(u.size());
    for (size_t k = 0; k < u.size(); k++) {
        const d3 p = cv.at(u[k]);
        out[k] = f3{(float)p.x, (float)p.y, (float)p.z};
    }
    return out;
}

// ---------------------------------------------------------------------------
// The host scene: what Sim(n) plus the user's edits hold before upload().
struct Curves {
    std::vector<std::vector<f3>> pts;
    std::vector<bool> closed;
    size_t total = 0;
    void add(std::vector<f3> c, bool isClosed) {
        total += c.size();
        pts.push_back(std::move(c));
        closed.push_back(isClosed);
    }
};

}  // namespace

struct yb_scene {
    int n = 0;
    float maxH = 1e-3f;
    yb_params p{};
    std::vector<yb_vertex> verts;
    std::vector<float> vels, qs, qRests, inv0;

    void init(int count) {
        n = count;
        verts.assign(n, yb_vertex{});
        vels.assign(3 * (size_t)n, 0.f);
        qs.assign(4 * (size_t)n, 0.f);
        qRests.assign(4 * (size_t)n, 0.f);
        inv0.assign(n, 0.f);
        if (yb_default_params(&p, n, maxH) != YB_OK ||
            yb_default_vertices(n, verts.data(), vels.data(), qs.data(), qRests.data()) != YB_OK)
            throw std::runtime_error("Too little vertices");
    }
};

namespace {

// createFromCurves (reader.cpp:13-48): concatenate, end each curve, and
// connect the two ends of a closed curve through connectionIndex.
std::unique_ptr<yb_scene> from_curves(const Curves& cv) {
    auto sc = std::make_unique<yb_scene>();
    sc->init((int)cv.total);
    size_t base = 0;
    for (size_t c = 0; c < cv.pts.size(); c++) {
        const auto& d = cv.pts[c];
        for (size_t k = 0; k < d.size(); k++) {
            yb_vertex& v = sc->verts[base + k];
            v.pos[0] = d[k].x;
            v.pos[1] = d[k].y;
            v.pos[2] = d[k].z;
        }
        sc->verts[base + d.size() - 1].flags = 0;
        if (cv.closed[c]) {
            sc->verts[base].connectionIndex = (int32_t)(base + d.size() - 1);
            sc->verts[base + d.size() - 1].connectionIndex = (int32_t)base;
        }
        base += d.size();
    }
    return sc;
}

#pragma pack(push, 1)
struct BccHeader {
    char sign[3];
    unsigned char byteCount;
    char curveType[2];
    char dimensions;
    char upDimension;
    uint64_t curveCount;
    uint64_t totalControlPointCount;
    char fileInfo[40];
};
#pragma pack(pop)
static_assert(sizeof(BccHeader) == 64, "BCC header is 64 bytes");

Curves read_bcc(const std::string& path, float segLen, const Affine& tf, bool breakUp,
                bool resample) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("Could not open file");
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
    BccHeader h;
    if (fread(&h, sizeof(h), 1, f) != 1 || h.sign[0] != 'B' || h.sign[1] != 'C' ||
        h.sign[2] != 'C' || h.byteCount != 0x44)
        throw std::runtime_error("Unsupported BCC file");
    const bool polyline = h.curveType[0] == 'P' && h.curveType[1] == 'L';
    if (!polyline && (h.curveType[0] != 'C' || h.curveType[1] != '0'))
        throw std::runtime_error("Not polyline or uniform CMR spline");
    if (h.dimensions != 3) throw std::runtime_error("Only curves in 3D are supported");
    Curves cv;
    for (uint64_t c = 0; c < h.curveCount; c++) {
        int32_t np = 0;
        if (fread(&np, sizeof(np), 1, f) != 1) throw std::runtime_error("Truncated BCC file");
        const bool isClosed = np < 0;
        np = np < 0 ? -np : np;
        std::vector<float> raw(3 * (size_t)np);
        if (np && fread(raw.data(), sizeof(float), raw.size(), f) != raw.size())
            throw std::runtime_error("Truncated BCC file");
        std::vector<f3> pts(np);
        for (int k = 0; k < np; k++) pts[k] = tf.apply(raw[3 * k], raw[3 * k + 1], raw[3 * k + 2]);
        if (resample && !polyline && np >= 4) pts = resample_cmr(pts, segLen);
        // the reference tests the control-point count, after resampling
        if (np < 3) continue;
        cv.add(std::move(pts), isClosed && !breakUp);
    }
    return cv;
}

Curves read_obj(const std::string& path, float segLen, const Affine& tf, bool breakUp,
                bool resample) {
    std::ifstream file(path);
    if (!file.is_open()) throw std::runtime_error("Could not open file.");
    std::vector<f3> vs;
    std::vector<std::vector<int>> lines;
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream iss(line);
        std::string tag;
        iss >> tag;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            iss >> x >> y >> z;
            vs.push_back(tf.apply(x, y, z));
        } else if (tag == "l") {
            std::vector<int> ids;
            int id;
            while (iss >> id) ids.push_back(id);
            if (!ids.empty()) lines.push_back(std::move(ids));
        }
    }
    Curves cv;
    for (const auto& l : lines) {
        std::vector<f3> c;
        bool isClosed = l.front() == l.back();
        for (int id : l) {
            if (id < 1 || id > (int)vs.size()) throw std::runtime_error("OBJ line index out of range");
            c.push_back(vs[id - 1]);
        }
        if (breakUp && isClosed) {
            c.pop_back();
            isClosed = false;
        }
        if (resample && c.size() >= 4) c = resample_cmr(c, segLen);
        if (c.size() < 4) continue;
        cv.add(std::move(c), isClosed);
    }
    return cv;
}

Curves read_poly(const std::string& path, float segLen, const Affine& tf, bool breakUp,
                 bool resample) {
    std::ifstream file(path);
    if (!file.is_open()) throw std::runtime_error("Could not open file.");
    std::vector<f3> pts;
    std::vector<std::pair<int, int>> segs;
    bool points = true;
    std::string line;
    while (std::getline(file, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "POINTS") {
            points = true;
            continue;
        }
        if (line == "POLYS") {
            points = false;
            continue;
        }
        if (line == "END") break;
        std::istringstream iss(line);
        int id;
        char colon;
        if (points) {
            float x, y, z;
            if (!(iss >> id >> colon >> x >> y >> z) || colon != ':')
                throw std::runtime_error("Error parsing .poly. Point parse error.");
            pts.push_back(tf.apply(x, y, z));
            if (id != (int)pts.size())
                throw std::runtime_error("Error parsing .poly. Point id mismatch.");
        } else {
            int a, b;
            if (!(iss >> id >> colon >> a >> b) || colon != ':')
                throw std::runtime_error("Error parsing .poly. Segment parse error.");
            if (a < 1 || b < 1 || a > (int)pts.size() || b > (int)pts.size())
                throw std::runtime_error("Error parsing .poly. Segment index out of range.");
            segs.emplace_back(a - 1, b - 1);
        }
    }
    // adjacency in listing order
    const int np = (int)pts.size();
    std::vector<std::vector<int>> adj(np);
    for (const auto& s : segs) {
        adj[s.first].push_back(s.second);
        adj[s.second].push_back(s.first);
    }
    Curves cv;
    std::vector<char> seen(np, 0);
    auto follow = [&](int start, std::vector<f3>& c, bool& returned) {
        int cur = start;
        returned = false;
        for (;;) {
            seen[cur] = 1;
            c.push_back(pts[cur]);
            int next = -1;
            for (int o : adj[cur]) {
                if (!seen[o]) {
                    next = o;
                    break;
                }
                if (o == start && c.size() > 2) returned = true;
            }
            if (next < 0) return cur;
            cur = next;
        }
    };
    // open curves: start at an end (a point with at most one neighbour)
    for (int i = 0; i < np; i++) {
        if (seen[i]) continue;
        if (adj[i].size() > 2) throw std::runtime_error("Graphs not allowed.");
        if (adj[i].size() > 1) continue;
        std::vector<f3> c;
        bool ret;
        follow(i, c, ret);
        if (c.size() < 4) continue;
        c = resample_cmr(c, segLen);  // open curves are always resampled
        cv.add(std::move(c), false);
    }
    // closed curves: everything left has exactly two neighbours
    for (int i = 0; i < np; i++) {
        if (seen[i]) continue;
        if (adj[i].size() != 2) throw std::runtime_error("Graphs not allowed.");
        std::vector<f3> c;
        bool ret;
        follow(i, c, ret);
        if (!ret) throw std::runtime_error("Closed curve not closed.");
        if (resample && c.size() >= 4) c = resample_cmr(c, segLen);
        if (c.size() < 4) continue;
        cv.add(std::move(c), !breakUp);
    }
    return cv;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t l = strlen(suf);
    return s.size() > l && s.compare(s.size() - l, l, suf) == 0;
}

Curves read_any(const std::string& path, float segLen, const Affine& tf, bool breakUp,
                bool resample) {
    if (ends_with(path, ".poly")) return read_poly(path, segLen, tf, breakUp, resample);
    if (ends_with(path, ".obj")) return read_obj(path, segLen, tf, breakUp, resample);
    return read_bcc(path, segLen, tf, breakUp, resample);
}

// ---------------------------------------------------------------------------
// glueEndpoints (jsonBuilder.cpp:174-212): every strand endpoint whose
// nearest vertex (index more than 2 away) lies within searchRadius is joined
// to it, if neither is joined yet; both move to their midpoint.
void glue_endpoints(int n, yb_vertex* v, float radius) {
    if (!(radius >= 0.f) || n <= 0) return;
    const float cell = radius > 0.f ? radius : 1.f;
    auto key = [&](const float* p) {
        const int64_t x = (int64_t)floorf(p[0] / cell), y = (int64_t)floorf(p[1] / cell),
                      z = (int64_t)floorf(p[2] / cell);
        return std::make_tuple(x, y, z);
    };
    struct H {
        size_t operator()(const std::tuple<int64_t, int64_t, int64_t>& k) const {
            return (size_t)(std::get<0>(k) * 73856093LL ^ std::get<1>(k) * 19349663LL ^
                            std::get<2>(k) * 83492791LL);
        }
    };
    std::unordered_map<std::tuple<int64_t, int64_t, int64_t>, std::vector<int>, H> grid;
    for (int j = 0; j < n; j++) grid[key(v[j].pos)].push_back(j);
    auto move_to = [&](int j, const float* np) {
        auto& bucket = grid[key(v[j].pos)];
        bucket.erase(std::find(bucket.begin(), bucket.end(), j));
        for (int k = 0; k < 3; k++) v[j].pos[k] = np[k];
        grid[key(v[j].pos)].push_back(j);
    };
    auto dist = [&](int a, int b) {
        const float dx = v[a].pos[0] - v[b].pos[0], dy = v[a].pos[1] - v[b].pos[1],
                    dz = v[a].pos[2] - v[b].pos[2];
        return sqrtf(dx * dx + dy * dy + dz * dz);
    };
    for (int i = 0; i < n; i++) {
        const bool hasPrev = i > 0 && (v[i - 1].flags & YB_FLAG_HAS_NEXT) != 0;
        const bool hasNext = (v[i].flags & YB_FLAG_HAS_NEXT) != 0;
        if (hasPrev == hasNext) continue;
        // the nearest vertex if it is within the radius (ties: smallest index)
        float best = INFINITY;
        int closest = -1;
        const auto k = key(v[i].pos);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    auto it = grid.find(std::make_tuple(std::get<0>(k) + dx, std::get<1>(k) + dy,
                                                        std::get<2>(k) + dz));
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        if (abs(i - j) <= 2) continue;
                        const float d = dist(i, j);
                        if (d < best || (d == best && j < closest)) {
                            best = d;
                            closest = j;
                        }
                    }
                }
        if (closest < 0 || !(best <= radius)) continue;
        if (v[closest].connectionIndex >= 0 || v[i].connectionIndex >= 0) continue;
        v[closest].connectionIndex = i;
        v[i].connectionIndex = closest;
        float mid[3];
        for (int c = 0; c < 3; c++) mid[c] = 0.5f * (v[i].pos[c] + v[closest].pos[c]);
        move_to(i, mid);
        move_to(closest, mid);
    }
}

// buildFromJSON (jsonBuilder.cpp:8-170).
std::unique_ptr<yb_scene> build_from_json(const std::string& path) {
    std::ifstream file(path, std::ifstream::binary);
    if (!file.is_open()) throw std::runtime_error("Error opening file");
    std::stringstream ss;
    ss << file.rdbuf();
    const std::string text = ss.str();
    const JVal root = JsonReader(text).parse();

    Affine tf;
    const JVal& t = member(root, "transform");
    if (!t.isNull())
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 4; j++) tf.m[i][j] = t.arr.at(i).arr.at(j).asFloat();

    std::string data = member(root, "curveFile").asString();
    // relative curve paths resolve against the working directory, as in the
    // reference; if that fails, against the JSON file's parent directory's parent
    // (configs/x.json naming configs/models/y.poly)
    auto exists = [](const std::string& p) { return (bool)std::ifstream(p); };
    if (!exists(data) && !data.empty() && data[0] != '/') {
        const size_t cut = path.find_last_of('/');
        const std::string dir = cut == std::string::npos ? "." : path.substr(0, cut);
        const size_t cut2 = dir.find_last_of('/');
        for (const std::string& base :
             {dir, cut2 == std::string::npos ? std::string(".") : dir.substr(0, cut2)})
            if (exists(base + "/" + data)) {
                data = base + "/" + data;
                break;
            }
    }
    const bool resample = !member(root, "resampleLength").isNull();
    const float segLen = resample ? member(root, "resampleLength").asFloat() : 0.f;
    const bool breakUp = member(root, "breakUpClosedCurves").isNull()
                             ? false
                             : member(root, "breakUpClosedCurves").asBool();
    std::unique_ptr<yb_scene> sc = from_curves(read_any(data, segLen, tf, breakUp, resample));
    yb_params& p = sc->p;

    if (!member(root, "curveRadius").isNull()) {
        const float r = member(root, "curveRadius").asFloat();
        constexpr float ratio = 0.05f;
        p.radius = ratio * r;
        p.barrierThickness = 2 * (1 - ratio) * r;
    }
    const JVal& sim = member(root, "simulation");
    double density = 1.0, kStretch = 5e5, kBend = 1e-1;
    if (!sim.isNull()) {
        auto num = [&](const char* k, float& dst) {
            if (!member(sim, k).isNull()) dst = member(sim, k).asFloat();
        };
        if (!member(sim, "density").isNull()) density = member(sim, "density").asFloat();
        num("frictionCoeff", p.frictionCoeff);
        if (!member(sim, "kStretch").isNull()) kStretch = member(sim, "kStretch").asFloat();
        if (!member(sim, "kBend").isNull()) kBend = member(sim, "kBend").asFloat();
        num("kCollision", p.kCollision);
        num("kFriction", p.kFriction);
        num("damping", p.damping);
        const JVal& g = member(sim, "gravity");
        if (!g.isNull())
            for (int k = 0; k < 3; k++) p.gravity[k] = g.arr.at(k).asFloat();
        num("drag", p.drag);
        if (!member(sim, "maxTimeStep").isNull()) sc->maxH = member(sim, "maxTimeStep").asFloat();
        if (!member(sim, "numIterations").isNull()) p.numItr = member(sim, "numIterations").asInt();
        if (!member(sim, "detectionPeriod").isNull())
            p.detectionPeriod = member(sim, "detectionPeriod").asInt();
        num("detectionScaler", p.detectionScaler);
        if (!member(sim, "stepLimit").isNull())
            p.useStepSizeLimit = member(sim, "stepLimit").asBool() ? 1 : 0;
        if (!member(sim, "velStepLimit").isNull())
            p.useVelocityRadius = member(sim, "velStepLimit").asBool() ? 1 : 0;
        num("bvhRebuildPeriod", p.bvhRebuildPeriod);
    }

    const int n = sc->n;
    int rc = yb_configure_host(n, sc->verts.data(), sc->qs.data(), sc->qRests.data(),
                               (float)density, sc->inv0.data(), &p);
    if (rc != YB_OK) throw std::runtime_error(yb_last_message());
    rc = yb_set_kstretch_host(n, sc->verts.data(), (float)kStretch);
    if (rc == YB_OK) rc = yb_set_kbend_host(n, sc->verts.data(), sc->qRests.data(), (float)kBend);
    if (rc != YB_OK) throw std::runtime_error(yb_last_message());

    if (!member(root, "glueEndpoints").isNull()) {
        const float r = member(root, "glueEndpoints").asFloat();
        if (r >= 0) glue_endpoints(n, sc->verts.data(), r);
    }

    const JVal& borders = member(root, "fixBorders");
    if (borders.isArray() && borders.arr.size() == 6) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], sc->verts[i].pos[k]);
                hi[k] = fmaxf(hi[k], sc->verts[i].pos[k]);
            }
        const float pad = 0.01f * p.radius;
        for (int k = 0; k < 3; k++) {
            lo[k] -= pad;
            hi[k] += pad;
        }
        for (int k = 0; k < 3; k++) {
            hi[k] -= borders.arr[2 * k].asFloat();
            lo[k] += borders.arr[2 * k + 1].asFloat();
        }
        for (int i = 0; i < n; i++) {
            const float* q = sc->verts[i].pos;
            bool inside = true;
            for (int k = 0; k < 3; k++) inside = inside && q[k] >= lo[k] && q[k] <= hi[k];
            if (!inside) sc->verts[i].invMass = 0;
        }
    }

    const JVal& fix = member(root, "fixVertex");
    if (fix.isArray())
        for (const JVal& e : fix.arr) {
            if (e.isArray()) {
                if (e.arr.size() < 3) continue;
                const float c[3] = {e.arr[0].asFloat(), e.arr[1].asFloat(), e.arr[2].asFloat()};
                auto d2 = [&](int j) {
                    const float dx = sc->verts[j].pos[0] - c[0], dy = sc->verts[j].pos[1] - c[1],
                                dz = sc->verts[j].pos[2] - c[2];
                    return dx * dx + dy * dy + dz * dz;
                };
                if (e.arr.size() >= 4) {
                    float r2 = e.arr[3].asFloat();
                    r2 *= r2;
                    for (int j = 0; j < n; j++)
                        if (d2(j) < r2) sc->verts[j].invMass = 0;
                } else {
                    float best = INFINITY;
                    int closest = -1;
                    for (int j = 0; j < n; j++)
                        if (d2(j) < best) {
                            best = d2(j);
                            closest = j;
                        }
                    if (closest >= 0) sc->verts[closest].invMass = 0;
                }
            } else {
                const int j = e.asInt();
                if (j < 0 || j >= n) throw std::runtime_error("fixVertex index out of range");
                sc->verts[j].invMass = 0;
            }
        }
    return sc;
}

int fail(const std::exception& e) {
    yb::set_message(e.what());
    return YB_ERR_VALIDATION;
}

}  // namespace

// ---- C ABI --------------------------------------------------------------------
extern "C" int yb_build_from_json(const char* path, yb_scene** out) {
    if (!path || !out) return YB_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        *out = build_from_json(path).release();
        return YB_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

extern "C" int yb_read_curves(const char* path, float targetSegLen, const float* transform,
                              int breakUpClosedCurves, int allowResample, yb_scene** out) {
    if (!path || !out) return YB_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        Affine tf;
        if (transform)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 4; j++) tf.m[i][j] = transform[4 * i + j];
        *out = from_curves(read_any(path, targetSegLen, tf, breakUpClosedCurves != 0,
                                    allowResample != 0))
                   .release();
        return YB_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

extern "C" int yb_scene_size(const yb_scene* s, int* numVerts) {
    if (!s || !numVerts) return YB_ERR_INVALID_ARGUMENT;
    *numVerts = s->n;
    return YB_OK;
}

extern "C" int yb_scene_get(const yb_scene* s, yb_params* p, float* maxH, yb_vertex* verts,
                            float* vels, float* qs, float* qRests, float* initialInvMasses) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)s->n;
    if (p) *p = s->p;
    if (maxH) *maxH = s->maxH;
    if (verts) memcpy(verts, s->verts.data(), sizeof(yb_vertex) * n);
    if (vels) memcpy(vels, s->vels.data(), sizeof(float) * 3 * n);
    if (qs) memcpy(qs, s->qs.data(), sizeof(float) * 4 * n);
    if (qRests) memcpy(qRests, s->qRests.data(), sizeof(float) * 4 * n);
    if (initialInvMasses) memcpy(initialInvMasses, s->inv0.data(), sizeof(float) * n);
    return YB_OK;
}

extern "C" void yb_scene_free(yb_scene* s) { delete s; }

extern "C" int yb_glue_endpoints(int n, yb_vertex* verts, float searchRadius) {
    if (n < 0 || (n && !verts)) return YB_ERR_INVALID_ARGUMENT;
    glue_endpoints(n, verts, searchRadius);
    return YB_OK;
}

// Runs of segments: a curve starts at a vertex with hasNext after one
// without, and ends at the first vertex without hasNext (fileExport.cpp).
static std::vector<std::pair<int, int>> curve_runs(int n, const yb_vertex* v) {
    std::vector<std::pair<int, int>> runs;
    bool last = false;
    int start = 0;
    for (int i = 0; i < n; i++) {
        const bool seg = (v[i].flags & YB_FLAG_HAS_NEXT) != 0;
        if (seg) {
            if (!last) start = i;
        } else if (last) {
            runs.emplace_back(start, i);
        }
        last = seg;
    }
    return runs;
}

extern "C" int yb_export_bcc(const char* path, int n, const yb_vertex* verts, int asPolyline) {
    if (!path || n < 0 || (n && !verts)) return YB_ERR_INVALID_ARGUMENT;
    FILE* f = fopen(path, "wb");
    if (!f) {
        yb::set_message("Error opening file");
        return YB_ERR_VALIDATION;
    }
    BccHeader h;
    memset(&h, 0, sizeof(h));
    h.sign[0] = 'B';
    h.sign[1] = 'C';
    h.sign[2] = 'C';
    h.byteCount = 0x44;
    h.curveType[0] = asPolyline ? 'P' : 'C';
    h.curveType[1] = asPolyline ? 'L' : '0';
    h.dimensions = 3;
    h.upDimension = 1;
    const auto runs = curve_runs(n, verts);
    h.curveCount = runs.size();
    h.totalControlPointCount = 0;
    for (const auto& r : runs) h.totalControlPointCount += (uint64_t)(r.second - r.first + 1);
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    for (const auto& r : runs) {
        const int32_t np = r.second - r.first + 1;
        ok = ok && fwrite(&np, sizeof(np), 1, f) == 1;
        for (int i = r.first; i <= r.second; i++) ok = ok && fwrite(verts[i].pos, 4, 3, f) == 3;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        yb::set_message("Error writing file");
        return YB_ERR_VALIDATION;
    }
    return YB_OK;
}

extern "C" int yb_export_obj(const char* path, int n, const yb_vertex* verts) {
    if (!path || n < 0 || (n && !verts)) return YB_ERR_INVALID_ARGUMENT;
    FILE* f = fopen(path, "w");
    if (!f) {
        yb::set_message("Error opening file");
        return YB_ERR_VALIDATION;
    }
    fprintf(f, "# YarnBall Sim\n\n");
    fprintf(f, "o YarnBall\n\n");
    fprintf(f, "# Vertices (in meters)\n");
    for (int i = 0; i < n; i++)
        fprintf(f, "v %.16f %.16f %.16f\n", verts[i].pos[0], verts[i].pos[1], verts[i].pos[2]);
    fprintf(f, "\n# Curves\n");
    bool last = false;
    for (int i = 0; i < n; i++) {
        const bool seg = (verts[i].flags & YB_FLAG_HAS_NEXT) != 0;
        if (seg) {
            if (!last) fprintf(f, "l %d", i + 1);
            fprintf(f, " %d", i + 2);
        } else if (last) {
            fprintf(f, "\n");
        }
        last = seg;
    }
    if (fclose(f) != 0) {
        yb::set_message("Error writing file");
        return YB_ERR_VALIDATION;
    }
    return YB_OK;
}
