// model_io.cpp — bundle-adjustment problems on disk (vba_io.h): the
// reference's plain-text problem dump (pipelines/slam/src/ba_problem_io.rs)
// and COLMAP sparse models, text and binary (crates/io/src/colmap/mod.rs),
// read into one host-side model whose vba_problem view feeds vba_optimize,
// and written back from it after a solve.
//
// Parsing follows the reference's readers token for token: the same line
// and token rules, Rust's number grammar (a leading '+', "inf"/"NaN" in any
// case, no hex), the same binary record-size checks and messages. Writers
// produce the reference's bytes: Rust's shortest round-trip `{}` for the
// problem dump, `{:.12}` with trailing zeros trimmed for COLMAP text,
// little-endian records for COLMAP binary.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/vba_io.h"

namespace {

thread_local std::string g_io_error;

struct IoError {
    int32_t code;
    std::string msg;
};

[[noreturn]] void fail(int32_t code, std::string msg) { throw IoError{code, std::move(msg)}; }

// ---- numbers as Rust's FromStr reads them ---------------------------------

bool parse_u64(const std::string& s, uint64_t* out) {
    size_t i = 0;
    if (i < s.size() && s[i] == '+') ++i;
    if (i == s.size()) return false;
    uint64_t v = 0;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        const uint64_t d = (uint64_t)(s[i] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

bool parse_i64(const std::string& s, int64_t* out) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return false;
    uint64_t v = 0;
    const uint64_t lim = neg ? (uint64_t)INT64_MAX + 1 : (uint64_t)INT64_MAX;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        const uint64_t d = (uint64_t)(s[i] - '0');
        if (v > (lim - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = neg ? (int64_t)(0 - v) : (int64_t)v;
    return true;
}

bool ieq(const std::string& a, size_t from, const char* b) {
    const size_t n = std::strlen(b);
    if (a.size() - from != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[from + i]) != b[i]) return false;
    return true;
}

// f64::from_str: [+-]? (inf | infinity | nan | digits [. digits] | . digits)
// ([eE] [+-]? digits)?, correctly rounded (glibc strtod is).
bool parse_f64(const std::string& s, double* out) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (ieq(s, i, "inf") || ieq(s, i, "infinity")) {
        *out = neg ? -INFINITY : INFINITY;
        return true;
    }
    if (ieq(s, i, "nan")) {
        *out = NAN;
        return true;
    }
    size_t digits = 0;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i, ++digits;
    if (i < s.size() && s[i] == '.') {
        ++i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
        size_t ed = 0;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i, ++ed;
        if (ed == 0) return false;
    }
    if (i != s.size()) return false;
    errno = 0;
    *out = std::strtod(s.c_str(), nullptr);  // ERANGE still gives the rounded value (0 or inf)
    return true;
}

// Rust's `{}` for f64: the shortest digits that read back to the same value,
// never in exponent form.
std::string display_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char buf[512];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::fixed);
    return std::string(buf, r.ptr);
}

// format_f64 (colmap/mod.rs:1371-1377): `{:.12}` with trailing zeros, then a
// trailing '.', trimmed.
std::string colmap_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.12f", v);
    std::string s(buf);
    while (!s.empty() && s.back() == '0') s.pop_back();
    if (!s.empty() && s.back() == '.') s.pop_back();
    return s;
}

std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        size_t j = i;
        while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
        if (j > i) t.emplace_back(line.substr(i, j - i));
        i = j;
    }
    return t;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// str::lines: split at '\n', a trailing "\r" dropped from each line, no final
// empty line after a trailing '\n'.
std::vector<std::string> rust_lines(const std::string& text) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < text.size()) {
        size_t j = text.find('\n', i);
        const bool nl = j != std::string::npos;
        if (!nl) j = text.size();
        std::string l = text.substr(i, j - i);
        if (nl && !l.empty() && l.back() == '\r') l.pop_back();
        out.push_back(std::move(l));
        i = nl ? j + 1 : j;
    }
    return out;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(VBA_ERR_IO, "I/O error: cannot open " + path + ": " + std::strerror(errno));
    std::ostringstream ss;
    ss << f.rdbuf();
    if (f.bad()) fail(VBA_ERR_IO, "I/O error: cannot read " + path);
    return ss.str();
}

void write_file(const std::string& path, const std::string& bytes) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) fail(VBA_ERR_IO, "I/O error: cannot create " + path + ": " + std::strerror(errno));
    f.write(bytes.data(), (std::streamsize)bytes.size());
    if (!f) fail(VBA_ERR_IO, "I/O error: cannot write " + path);
}

// fs::create_dir_all
void create_dir_all(const std::string& dir) {
    std::string cur;
    size_t i = 0;
    while (i <= dir.size()) {
        size_t j = dir.find('/', i);
        if (j == std::string::npos) j = dir.size();
        cur = dir.substr(0, j);
        if (!cur.empty()) {
            struct stat st;
            if (stat(cur.c_str(), &st) != 0) {
                if (mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
                    fail(VBA_ERR_IO, "I/O error: cannot create directory " + cur + ": " + std::strerror(errno));
            } else if (!S_ISDIR(st.st_mode)) {
                fail(VBA_ERR_IO, "I/O error: " + cur + " is not a directory");
            }
        }
        i = j + 1;
    }
}

std::string join(const std::string& dir, const char* name) {
    return dir.empty() || dir.back() == '/' ? dir + name : dir + "/" + name;
}

// UnitQuaternion::from_quaternion: every coordinate divided by the norm, the
// squared norm summed over nalgebra's coordinate order (i, j, k, w).
void normalize_quat(double q[4]) {
    const double n = std::sqrt(((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]) + q[0] * q[0]);
    for (int c = 0; c < 4; ++c) q[c] /= n;
}

// ---- COLMAP camera models (colmap/mod.rs:1297-1500; camera.rs:44-59) -----

struct ModelInfo {
    const char* name;
    int32_t colmap_id;
    int params;
    int32_t vba;  // enum vba_camera_model, -1 when the solver has none
};
const ModelInfo kModels[] = {
    {"SIMPLE_PINHOLE", 0, 3, VBA_CAMERA_SIMPLE_PINHOLE},
    {"PINHOLE", 1, 4, VBA_CAMERA_PINHOLE},
    {"SIMPLE_RADIAL", 2, 4, VBA_CAMERA_SIMPLE_RADIAL},
    {"RADIAL", 3, 5, VBA_CAMERA_RADIAL},
    {"OPENCV", 4, 8, VBA_CAMERA_OPENCV},
    {"OPENCV_FISHEYE", 5, 8, -1},
    {"FULL_OPENCV", 6, 12, -1},
    {"FOV", 7, 5, -1},
    {"SIMPLE_RADIAL_FISHEYE", 8, 4, -1},
    {"RADIAL_FISHEYE", 9, 5, -1},
    {"THIN_PRISM_FISHEYE", 10, 12, -1},
    {"EQUIRECTANGULAR", 17, 2, -1},
};

const ModelInfo* model_by_name(const std::string& n) {
    for (const auto& m : kModels)
        if (n == m.name) return &m;
    return nullptr;
}
const ModelInfo* model_by_id(int32_t id) {
    for (const auto& m : kModels)
        if (m.colmap_id == id) return &m;
    return nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// The model: the problem at the level of `BundleAdjustment` (poses,
// landmarks, the observation list) plus what a COLMAP model adds (cameras,
// each image's camera and its full keypoint list).
// ---------------------------------------------------------------------------
struct Cam {
    uint64_t id;
    std::string model;  // COLMAP name
    uint32_t width, height;
    std::vector<double> params;
};
struct Keypoint {
    double x, y;
    bool observed;  // POINT3D_ID >= 0
    uint64_t point; // the POINT3D_ID when observed
};

struct vba_model {
    std::vector<Cam> cams;  // ascending id
    std::vector<uint64_t> pose_ids;  // ascending
    std::vector<double> poses;       // 7 per pose
    std::vector<uint8_t> pose_fixed;
    std::vector<uint64_t> pose_cam;  // camera id per pose
    bool has_keypoints = false;      // keypoints below hold the images' POINTS2D[]
    std::vector<std::vector<Keypoint>> keypoints;
    std::vector<uint64_t> lm_ids;  // ascending
    std::vector<double> lms;       // 3 per landmark
    std::vector<uint64_t> obs_pose, obs_lm;
    std::vector<double> obs_xy;
    std::vector<uint32_t> obs_kp;  // keypoint index per observation (with has_keypoints)
    bool weighted = false;
    std::vector<double> weights;

    // the view (vba_model_problem)
    vba_problem view{};
    std::vector<uint64_t> rig_pose, rig_lm;
    std::vector<double> rig_xy;
    std::vector<uint32_t> rig_sensor;
    std::vector<vba_sensor> sensors;
};

namespace {

size_t pose_index(const vba_model& m, uint64_t id) {
    auto it = std::lower_bound(m.pose_ids.begin(), m.pose_ids.end(), id);
    return it != m.pose_ids.end() && *it == id ? (size_t)(it - m.pose_ids.begin()) : SIZE_MAX;
}

const Cam* find_cam(const vba_model& m, uint64_t id) {
    for (const auto& c : m.cams)
        if (c.id == id) return &c;
    return nullptr;
}

// ---- ba_problem_io.rs ------------------------------------------------------

vba_model* read_ba_problem(const std::string& path) {
    const std::vector<std::string> lines = rust_lines(read_file(path));
    size_t li = 0;
    auto next = [&](const char* tag) {
        if (li >= lines.size()) fail(VBA_ERR_FORMAT, std::string("invalid data: ") + tag);
        return split_ws(lines[li++]);
    };
    auto f = [&](const std::vector<std::string>& v, size_t i) {
        double x;
        if (i >= v.size() || !parse_f64(v[i], &x)) fail(VBA_ERR_FORMAT, "invalid data: float");
        return x;
    };
    auto u = [&](const std::vector<std::string>& v, size_t i) {
        uint64_t x;
        if (i >= v.size() || !parse_u64(v[i], &x)) fail(VBA_ERR_FORMAT, "invalid data: int");
        return x;
    };
    auto* m = new vba_model();
    try {
        const auto cam = next("camera");
        if (cam.size() != 7 || cam[0] != "camera") fail(VBA_ERR_FORMAT, "invalid data: camera line");
        Cam c{1, "PINHOLE", (uint32_t)u(cam, 1), (uint32_t)u(cam, 2), {f(cam, 3), f(cam, 4), f(cam, 5), f(cam, 6)}};
        m->cams.push_back(c);
        // BTreeMap inserts: a repeated id keeps its last value; fix_pose only
        // ever adds to the fixed set
        std::map<uint64_t, std::vector<double>> poses;
        std::set<uint64_t> fixed;
        const uint64_t np = u(next("poses"), 1);
        for (uint64_t i = 0; i < np; ++i) {
            const auto v = next("pose");
            const uint64_t id = u(v, 0);
            double q[4] = {f(v, 1), f(v, 2), f(v, 3), f(v, 4)};
            normalize_quat(q);
            poses[id] = {q[0], q[1], q[2], q[3], f(v, 5), f(v, 6), f(v, 7)};
            if (v.size() < 9) fail(VBA_ERR_FORMAT, "invalid data: pose");
            if (v[8] == "1") fixed.insert(id);
        }
        std::map<uint64_t, std::vector<double>> lms;
        const uint64_t nl = u(next("landmarks"), 1);
        for (uint64_t i = 0; i < nl; ++i) {
            const auto v = next("landmark");
            const uint64_t id = u(v, 0);
            lms[id] = {f(v, 1), f(v, 2), f(v, 3)};
        }
        const auto head = next("observations");
        const uint64_t no = u(head, 1);
        m->weighted = head.size() > 2 && head[2] == "1";
        for (uint64_t i = 0; i < no; ++i) {
            const auto v = next("observation");
            m->obs_pose.push_back(u(v, 0));
            m->obs_lm.push_back(u(v, 1));
            m->obs_xy.push_back(f(v, 2));
            m->obs_xy.push_back(f(v, 3));
            if (m->weighted) m->weights.push_back(f(v, 4));
        }
        for (const auto& [id, p] : poses) {
            m->pose_ids.push_back(id);
            m->poses.insert(m->poses.end(), p.begin(), p.end());
            m->pose_fixed.push_back(fixed.count(id) ? 1 : 0);
            m->pose_cam.push_back(1);
        }
        for (const auto& [id, p] : lms) {
            m->lm_ids.push_back(id);
            m->lms.insert(m->lms.end(), p.begin(), p.end());
        }
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

void write_ba_problem(const vba_model& m, const std::string& path) {
    if (m.cams.size() != 1 || m.cams[0].model != "PINHOLE" || m.cams[0].params.size() < 4)
        fail(VBA_ERR_EXPORT, "the problem dump holds one PINHOLE camera");
    const Cam& c = m.cams[0];
    std::string s;
    s.reserve(64 * (m.pose_ids.size() + m.lm_ids.size() + m.obs_pose.size()) + 64);
    s += "camera " + std::to_string(c.width) + " " + std::to_string(c.height);
    for (int i = 0; i < 4; ++i) s += " " + display_f64(c.params[i]);
    s += "\nposes " + std::to_string(m.pose_ids.size()) + "\n";
    for (size_t i = 0; i < m.pose_ids.size(); ++i) {
        s += std::to_string(m.pose_ids[i]);
        for (int k = 0; k < 7; ++k) s += " " + display_f64(m.poses[7 * i + k]);
        s += m.pose_fixed[i] ? " 1\n" : " 0\n";
    }
    s += "landmarks " + std::to_string(m.lm_ids.size()) + "\n";
    for (size_t i = 0; i < m.lm_ids.size(); ++i) {
        s += std::to_string(m.lm_ids[i]);
        for (int k = 0; k < 3; ++k) s += " " + display_f64(m.lms[3 * i + k]);
        s += "\n";
    }
    s += "observations " + std::to_string(m.obs_pose.size()) + (m.weighted ? " 1\n" : " 0\n");
    for (size_t i = 0; i < m.obs_pose.size(); ++i) {
        s += std::to_string(m.obs_pose[i]) + " " + std::to_string(m.obs_lm[i]) + " " + display_f64(m.obs_xy[2 * i]) +
             " " + display_f64(m.obs_xy[2 * i + 1]);
        if (m.weighted) s += " " + display_f64(m.weights[i]);
        s += "\n";
    }
    write_file(path, s);
}

// ---- COLMAP readers (colmap/mod.rs:148-196, 989-1295) ---------------------

struct Image {
    uint64_t id, camera_id;
    double pose[7];
    std::vector<Keypoint> kps;
};
struct Point {
    uint64_t id;
    double xyz[3];
};

[[noreturn]] void bad_line(const char* file, const std::string& line) {
    fail(VBA_ERR_FORMAT, std::string("invalid COLMAP line in ") + file + ": " + line);
}

uint64_t tok_u64(const std::string& t) {
    uint64_t v;
    if (!parse_u64(t, &v)) fail(VBA_ERR_FORMAT, "integer parse error: " + t);
    return v;
}
uint32_t tok_u32(const std::string& t) {
    uint64_t v;
    if (!parse_u64(t, &v) || v > UINT32_MAX) fail(VBA_ERR_FORMAT, "integer parse error: " + t);
    return (uint32_t)v;
}
int64_t tok_i64(const std::string& t) {
    int64_t v;
    if (!parse_i64(t, &v)) fail(VBA_ERR_FORMAT, "integer parse error: " + t);
    return v;
}
double tok_f64(const std::string& t) {
    double v;
    if (!parse_f64(t, &v)) fail(VBA_ERR_FORMAT, "float parse error: " + t);
    return v;
}

std::vector<Cam> parse_cameras_txt(const std::string& text) {
    std::vector<Cam> out;
    for (const auto& raw : rust_lines(text)) {
        const std::string line = trim(raw);
        if (line.empty() || line[0] == '#') continue;
        const auto t = split_ws(line);
        if (t.size() < 5) bad_line("cameras.txt", line);
        Cam c{tok_u64(t[0]), t[1], tok_u32(t[2]), tok_u32(t[3]), {}};
        for (size_t i = 4; i < t.size(); ++i) c.params.push_back(tok_f64(t[i]));
        out.push_back(std::move(c));
    }
    return out;
}

std::vector<Point> parse_points3d_txt(const std::string& text) {
    std::vector<Point> out;
    for (const auto& raw : rust_lines(text)) {
        const std::string line = trim(raw);
        if (line.empty() || line[0] == '#') continue;
        const auto t = split_ws(line);
        if (t.size() < 8) bad_line("points3D.txt", line);
        out.push_back(Point{tok_u64(t[0]), {tok_f64(t[1]), tok_f64(t[2]), tok_f64(t[3])}});
    }
    return out;
}

std::vector<Image> parse_images_txt(const std::string& text) {
    std::vector<std::string> lines;
    for (const auto& raw : rust_lines(text)) {
        std::string l = trim(raw);
        if (!l.empty() && l[0] == '#') continue;
        lines.push_back(std::move(l));
    }
    std::vector<Image> out;
    size_t i = 0;
    while (i < lines.size()) {
        const std::string header = lines[i++];
        if (header.empty()) continue;
        const auto t = split_ws(header);
        if (t.size() < 10) bad_line("images.txt", header);
        Image im;
        im.id = tok_u64(t[0]);
        double q[4] = {tok_f64(t[1]), tok_f64(t[2]), tok_f64(t[3]), tok_f64(t[4])};
        const double tx = tok_f64(t[5]), ty = tok_f64(t[6]), tz = tok_f64(t[7]);
        im.camera_id = tok_u64(t[8]);
        normalize_quat(q);
        const double pose[7] = {q[0], q[1], q[2], q[3], tx, ty, tz};
        std::copy(pose, pose + 7, im.pose);
        const std::string points = i < lines.size() ? lines[i++] : std::string();
        const auto p = split_ws(points);
        if (p.size() % 3 != 0) bad_line("images.txt", points);
        for (size_t k = 0; k < p.size(); k += 3) {
            const double x = tok_f64(p[k]), y = tok_f64(p[k + 1]);
            const int64_t id = tok_i64(p[k + 2]);
            im.kps.push_back(Keypoint{x, y, id >= 0, (uint64_t)id});
        }
        out.push_back(std::move(im));
    }
    return out;
}

// BinaryReader (colmap/mod.rs:1604-1740).
struct BinReader {
    const char* file;
    const std::string& b;
    size_t off = 0;
    [[noreturn]] void bad(const std::string& msg) const {
        fail(VBA_ERR_FORMAT, std::string("invalid COLMAP binary in ") + file + ": " + msg);
    }
    const char* take(size_t n) {
        if (off + n > b.size()) bad("unexpected end of file at byte " + std::to_string(off));
        const char* p = b.data() + off;
        off += n;
        return p;
    }
    template <typename T>
    T get() {
        T v;
        std::memcpy(&v, take(sizeof(T)), sizeof(T));  // little-endian host
        return v;
    }
    size_t count(const char* field, size_t min_bytes) {
        const uint64_t n = get<uint64_t>();
        const size_t rem = b.size() - off;
        if (min_bytes && n > SIZE_MAX / min_bytes)
            bad(std::string(field) + "=" + std::to_string(n) + " overflows minimum record-byte calculation");
        const size_t need = (size_t)n * min_bytes;
        if (need > rem)
            bad(std::string(field) + "=" + std::to_string(n) + " requires at least " + std::to_string(need) +
                " remaining byte(s), but only " + std::to_string(rem) + " remain");
        return (size_t)n;
    }
    std::string cstring() {
        const size_t start = off;
        while (off < b.size() && b[off] != 0) ++off;
        if (off >= b.size()) bad("unterminated string");
        std::string s = b.substr(start, off - start);
        if (!valid_utf8(s)) bad("invalid UTF-8 string");
        ++off;
        return s;
    }
    void finish() const {
        if (off != b.size()) bad(std::to_string(b.size() - off) + " trailing byte(s) after parsing");
    }
    // str::from_utf8: well-formed UTF-8 (no overlong forms, no surrogates,
    // nothing above U+10FFFF)
    static bool valid_utf8(const std::string& s) {
        size_t i = 0;
        while (i < s.size()) {
            const unsigned char c = (unsigned char)s[i];
            int extra;
            uint32_t cp;
            if (c < 0x80) {
                ++i;
                continue;
            } else if ((c >> 5) == 6) {
                extra = 1, cp = c & 0x1f;
            } else if ((c >> 4) == 14) {
                extra = 2, cp = c & 0x0f;
            } else if ((c >> 3) == 30) {
                extra = 3, cp = c & 0x07;
            } else {
                return false;
            }
            if (i + (size_t)extra >= s.size()) return false;
            for (int k = 1; k <= extra; ++k) {
                const unsigned char d = (unsigned char)s[i + k];
                if ((d >> 6) != 2) return false;
                cp = (cp << 6) | (d & 0x3f);
            }
            if ((extra == 1 && cp < 0x80) || (extra == 2 && cp < 0x800) || (extra == 3 && cp < 0x10000) ||
                cp > 0x10ffff || (cp >= 0xd800 && cp <= 0xdfff))
                return false;
            i += (size_t)extra + 1;
        }
        return true;
    }
};

std::vector<Cam> parse_cameras_bin(const std::string& bytes) {
    BinReader r{"cameras.bin", bytes};
    const size_t n = r.count("camera_count", 4 + 4 + 8 + 8 + 2 * 8);
    std::vector<Cam> out;
    for (size_t i = 0; i < n; ++i) {
        Cam c;
        c.id = r.get<uint32_t>();
        const int32_t model_id = r.get<int32_t>();
        const uint64_t w = r.get<uint64_t>(), h = r.get<uint64_t>();
        if (w > UINT32_MAX) r.bad("camera width does not fit in u32");
        if (h > UINT32_MAX) r.bad("camera height does not fit in u32");
        c.width = (uint32_t)w;
        c.height = (uint32_t)h;
        const ModelInfo* mi = model_by_id(model_id);
        if (!mi) r.bad("unsupported camera model id " + std::to_string(model_id));
        c.model = mi->name;
        for (int k = 0; k < mi->params; ++k) c.params.push_back(r.get<double>());
        out.push_back(std::move(c));
    }
    r.finish();
    return out;
}

std::vector<Image> parse_images_bin(const std::string& bytes) {
    BinReader r{"images.bin", bytes};
    const size_t n = r.count("image_count", 4 + 7 * 8 + 4 + 1 + 8);
    std::vector<Image> out;
    for (size_t i = 0; i < n; ++i) {
        Image im;
        im.id = r.get<uint32_t>();
        double q[4];
        for (double& v : q) v = r.get<double>();
        double t[3];
        for (double& v : t) v = r.get<double>();
        im.camera_id = r.get<uint32_t>();
        r.cstring();
        normalize_quat(q);
        const double pose[7] = {q[0], q[1], q[2], q[3], t[0], t[1], t[2]};
        std::copy(pose, pose + 7, im.pose);
        const size_t np = r.count("point_count", 3 * 8);
        for (size_t k = 0; k < np; ++k) {
            const double x = r.get<double>(), y = r.get<double>();
            const int64_t id = r.get<int64_t>();
            im.kps.push_back(Keypoint{x, y, id >= 0, (uint64_t)id});
        }
        out.push_back(std::move(im));
    }
    r.finish();
    return out;
}

std::vector<Point> parse_points3d_bin(const std::string& bytes) {
    BinReader r{"points3D.bin", bytes};
    const size_t n = r.count("point3D_count", 8 + 3 * 8 + 3 + 8 + 8);
    std::vector<Point> out;
    for (size_t i = 0; i < n; ++i) {
        Point p;
        p.id = r.get<uint64_t>();
        for (double& v : p.xyz) v = r.get<double>();
        r.take(3);            // RGB
        r.get<double>();      // ERROR
        const size_t tl = r.count("track_length", 2 * 4);
        r.take(8 * tl);       // (IMAGE_ID, POINT2D_IDX) pairs
        out.push_back(p);
    }
    r.finish();
    return out;
}

// VisualMap from the three record lists: BTreeMaps keyed by id, a repeated
// id keeping its last record. Observations are the keypoints with a
// POINT3D_ID >= 0 (mod.rs:1088-1098), by image id then keypoint index.
vba_model* model_from_colmap(std::vector<Cam> cams, std::vector<Image> images, std::vector<Point> points) {
    std::map<uint64_t, Cam> cm;
    for (auto& c : cams) cm[c.id] = std::move(c);
    std::map<uint64_t, Image> im;
    for (auto& i : images) im[i.id] = std::move(i);
    std::map<uint64_t, Point> pm;
    for (auto& p : points) pm[p.id] = p;
    auto* m = new vba_model();
    for (auto& [id, c] : cm) m->cams.push_back(std::move(c));
    m->has_keypoints = true;
    for (auto& [id, i] : im) {
        m->pose_ids.push_back(id);
        m->poses.insert(m->poses.end(), i.pose, i.pose + 7);
        m->pose_fixed.push_back(0);
        m->pose_cam.push_back(i.camera_id);
        for (size_t k = 0; k < i.kps.size(); ++k)
            if (i.kps[k].observed) {
                m->obs_pose.push_back(id);
                m->obs_lm.push_back(i.kps[k].point);
                m->obs_xy.push_back(i.kps[k].x);
                m->obs_xy.push_back(i.kps[k].y);
                m->obs_kp.push_back((uint32_t)k);
            }
        m->keypoints.push_back(std::move(i.kps));
    }
    for (auto& [id, p] : pm) {
        m->lm_ids.push_back(id);
        m->lms.insert(m->lms.end(), p.xyz, p.xyz + 3);
    }
    // The SfM's global-BA gauge (refinement.rs:316-352): the first pose and
    // the one whose camera centre is farthest from it (the first such on a
    // tie) are fixed.
    if (!m->pose_ids.empty()) {
        m->pose_fixed[0] = 1;
        auto centre = [&](size_t i, double c[3]) {
            const double* p = &m->poses[7 * i];
            const double w = p[0], x = -p[1], y = -p[2], z = -p[3];  // the inverse rotation
            const double v[3] = {-p[4], -p[5], -p[6]};
            // q v q^* for a unit quaternion, as R(q) v
            const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
                                 2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                                 2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
            for (int r = 0; r < 3; ++r) c[r] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
        };
        double a[3];
        centre(0, a);
        size_t far = SIZE_MAX;
        double best = 0.0;
        for (size_t i = 1; i < m->pose_ids.size(); ++i) {
            double c[3];
            centre(i, c);
            const double d2 = (c[0] - a[0]) * (c[0] - a[0]) + (c[1] - a[1]) * (c[1] - a[1]) + (c[2] - a[2]) * (c[2] - a[2]);
            if (d2 > best) {
                best = d2;
                far = i;
            }
        }
        if (far != SIZE_MAX) m->pose_fixed[far] = 1;
    }
    return m;
}

// ---- COLMAP writers (colmap/mod.rs:198-330, 900-987, 1297-1500) ----------

struct Record {
    const ModelInfo* info;
    std::vector<double> params;
};

// validated_colmap_camera_record: colmap_camera_record's rewrites (a
// PINHOLE with an all-zero distortion tail is a PINHOLE; a PINHOLE or OPENCV
// with 5..7 parameters is an OPENCV padded with zeros), then a known model
// with exactly its parameter count.
Record camera_record(const Cam& c) {
    std::string model = c.model;
    std::vector<double> p = c.params;
    const bool zero_tail = p.size() > 4 && std::all_of(p.begin() + 4, p.end(), [](double k) { return k == 0.0; });
    if (model == "PINHOLE" && zero_tail) {
        p.resize(4);
    } else if ((model == "PINHOLE" || model == "OPENCV") && p.size() > 4 && p.size() < 8) {
        model = "OPENCV";
        p.resize(8, 0.0);
    }
    const ModelInfo* mi = model_by_name(model);
    if (!mi)
        fail(VBA_ERR_EXPORT, "invalid COLMAP export input: cannot encode camera " + std::to_string(c.id) + " model " +
                                 model + " as a COLMAP model");
    if ((int)p.size() != mi->params)
        fail(VBA_ERR_EXPORT, "invalid COLMAP export input: camera " + std::to_string(c.id) + " model " + model +
                                 " needs " + std::to_string(mi->params) + " params for COLMAP, got " +
                                 std::to_string(p.size()));
    return Record{mi, std::move(p)};
}

// Each image's POINTS2D[] as written: the model's own keypoint list, or the
// image's observations in list order when the model has none.
std::vector<std::vector<Keypoint>> write_keypoints(const vba_model& m, std::vector<uint32_t>* obs_kp) {
    if (m.has_keypoints) {
        *obs_kp = m.obs_kp;
        return m.keypoints;
    }
    std::vector<std::vector<Keypoint>> kps(m.pose_ids.size());
    obs_kp->assign(m.obs_pose.size(), 0);
    for (size_t i = 0; i < m.obs_pose.size(); ++i) {
        const size_t p = pose_index(m, m.obs_pose[i]);
        if (p == SIZE_MAX) continue;
        (*obs_kp)[i] = (uint32_t)kps[p].size();
        kps[p].push_back(Keypoint{m.obs_xy[2 * i], m.obs_xy[2 * i + 1], true, m.obs_lm[i]});
    }
    return kps;
}

// The landmarks' tracks: every observation of the landmark as (image id,
// keypoint index), sorted.
std::vector<std::vector<std::pair<uint64_t, uint64_t>>> tracks(const vba_model& m, const std::vector<uint32_t>& kp) {
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> t(m.lm_ids.size());
    for (size_t i = 0; i < m.obs_lm.size(); ++i) {
        auto it = std::lower_bound(m.lm_ids.begin(), m.lm_ids.end(), m.obs_lm[i]);
        if (it == m.lm_ids.end() || *it != m.obs_lm[i] || pose_index(m, m.obs_pose[i]) == SIZE_MAX) continue;
        t[it - m.lm_ids.begin()].emplace_back(m.obs_pose[i], kp[i]);
    }
    for (auto& v : t) std::sort(v.begin(), v.end());
    return t;
}

// A keypoint's POINT3D_ID in images.bin (format_images_bin, mod.rs:284-296).
int64_t binary_point_id(const Keypoint& k) {
    if (!k.observed) return -1;
    if (k.point > (uint64_t)INT64_MAX)
        fail(VBA_ERR_EXPORT, "invalid COLMAP export input: landmark id " + std::to_string(k.point) +
                                 " does not fit in COLMAP's signed POINT3D_ID");
    return (int64_t)k.point;
}

uint32_t u32_id(uint64_t v, const char* what) {
    if (v > UINT32_MAX)
        fail(VBA_ERR_EXPORT, std::string("invalid COLMAP export input: ") + what + " " + std::to_string(v) +
                                 " does not fit in u32 (COLMAP binary models use 32-bit ids)");
    return (uint32_t)v;
}

template <typename T>
void put(std::string& b, T v) {
    char tmp[sizeof(T)];
    std::memcpy(tmp, &v, sizeof(T));
    b.append(tmp, sizeof(T));
}

void write_colmap(const vba_model& m, const std::string& dir, bool binary) {
    std::vector<Record> recs;
    for (const auto& c : m.cams) recs.push_back(camera_record(c));
    std::vector<uint32_t> kp_of_obs;
    const auto kps = write_keypoints(m, &kp_of_obs);
    const auto trk = tracks(m, kp_of_obs);
    std::string cams, imgs, pts;
    if (!binary) {
        cams = "# CAMERA_ID MODEL WIDTH HEIGHT PARAMS[]\n";
        for (size_t i = 0; i < m.cams.size(); ++i) {
            cams += std::to_string(m.cams[i].id) + " " + recs[i].info->name + " " + std::to_string(m.cams[i].width) +
                    " " + std::to_string(m.cams[i].height);
            for (double p : recs[i].params) cams += " " + colmap_f64(p);
            cams += "\n";
        }
        imgs = "# IMAGE_ID QW QX QY QZ TX TY TZ CAMERA_ID NAME\n# POINTS2D[] as X Y POINT3D_ID\n";
        for (size_t i = 0; i < m.pose_ids.size(); ++i) {
            imgs += std::to_string(m.pose_ids[i]);
            for (int k = 0; k < 7; ++k) imgs += " " + colmap_f64(m.poses[7 * i + k]);
            imgs += " " + std::to_string(m.pose_cam[i]) + " image_" + std::to_string(m.pose_ids[i]) + ".jpg\n";
            for (size_t k = 0; k < kps[i].size(); ++k) {
                if (k) imgs += " ";
                imgs += colmap_f64(kps[i][k].x) + " " + colmap_f64(kps[i][k].y) + " " +
                        (kps[i][k].observed ? std::to_string(kps[i][k].point) : std::string("-1"));
            }
            imgs += "\n";
        }
        pts = "# POINT3D_ID X Y Z R G B ERROR TRACK[] as IMAGE_ID POINT2D_IDX\n";
        for (size_t i = 0; i < m.lm_ids.size(); ++i) {
            pts += std::to_string(m.lm_ids[i]);
            for (int k = 0; k < 3; ++k) pts += " " + colmap_f64(m.lms[3 * i + k]);
            pts += " 255 255 255 0";
            for (const auto& [img, kp] : trk[i]) pts += " " + std::to_string(img) + " " + std::to_string(kp);
            pts += "\n";
        }
    } else {
        put<uint64_t>(cams, m.cams.size());
        for (size_t i = 0; i < m.cams.size(); ++i) {
            put<uint32_t>(cams, u32_id(m.cams[i].id, "camera id"));
            put<int32_t>(cams, recs[i].info->colmap_id);
            put<uint64_t>(cams, m.cams[i].width);
            put<uint64_t>(cams, m.cams[i].height);
            for (double p : recs[i].params) put<double>(cams, p);
        }
        put<uint64_t>(imgs, m.pose_ids.size());
        for (size_t i = 0; i < m.pose_ids.size(); ++i) {
            put<uint32_t>(imgs, u32_id(m.pose_ids[i], "image id"));
            for (int k = 0; k < 7; ++k) put<double>(imgs, m.poses[7 * i + k]);
            put<uint32_t>(imgs, u32_id(m.pose_cam[i], "camera id"));
            imgs += "image_" + std::to_string(m.pose_ids[i]) + ".jpg";
            imgs.push_back('\0');
            put<uint64_t>(imgs, kps[i].size());
            for (const auto& k : kps[i]) {
                put<double>(imgs, k.x);
                put<double>(imgs, k.y);
                put<int64_t>(imgs, binary_point_id(k));
            }
        }
        put<uint64_t>(pts, m.lm_ids.size());
        for (size_t i = 0; i < m.lm_ids.size(); ++i) {
            put<uint64_t>(pts, m.lm_ids[i]);
            for (int k = 0; k < 3; ++k) put<double>(pts, m.lms[3 * i + k]);
            pts.append("\xff\xff\xff", 3);
            put<double>(pts, 0.0);
            put<uint64_t>(pts, trk[i].size());
            for (const auto& [img, kp] : trk[i]) {
                put<uint32_t>(pts, u32_id(img, "track image id"));
                put<uint32_t>(pts, u32_id(kp, "track POINT2D_IDX"));
            }
        }
    }
    create_dir_all(dir);
    write_file(join(dir, binary ? "cameras.bin" : "cameras.txt"), cams);
    write_file(join(dir, binary ? "images.bin" : "images.txt"), imgs);
    write_file(join(dir, binary ? "points3D.bin" : "points3D.txt"), pts);
}

// ---- the vba_problem view ---------------------------------------------------

vba_camera to_vba_camera(const Cam& c) {
    const ModelInfo* mi = model_by_name(c.model);
    if (!mi || mi->vba < 0)
        fail(VBA_ERR_UNSUPPORTED_CAMERA_MODEL, "camera " + std::to_string(c.id) + " model " + c.model +
                                                   " is not one the solver supports");
    // PINHOLE may carry the reference's two trailing radial slots
    const size_t n = c.params.size();
    const bool ok = (int)n == mi->params || (mi->vba == VBA_CAMERA_PINHOLE && n == 6);
    if (!ok)
        fail(VBA_ERR_UNSUPPORTED_CAMERA_MODEL, "camera " + std::to_string(c.id) + " model " + c.model + " has " +
                                                   std::to_string(n) + " params");
    vba_camera v{};
    v.model = mi->vba;
    v.width = c.width;
    v.height = c.height;
    v.num_params = (uint32_t)n;
    for (size_t i = 0; i < n; ++i) v.params[i] = c.params[i];
    return v;
}

void build_view(vba_model& m) {
    // cameras that observations use, ascending id
    std::vector<uint64_t> used;
    for (size_t i = 0; i < m.obs_pose.size(); ++i) {
        const size_t p = pose_index(m, m.obs_pose[i]);
        if (p != SIZE_MAX) used.push_back(m.pose_cam[p]);
    }
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    for (uint64_t id : used)
        if (!find_cam(m, id))
            fail(VBA_ERR_INVALID_ARGUMENT, "an image uses camera " + std::to_string(id) + ", which the model does not have");
    vba_problem& v = m.view;
    v = vba_problem{};
    const bool rig = used.size() > 1;
    if (!m.cams.empty()) v.camera = to_vba_camera(used.empty() ? m.cams[0] : *find_cam(m, used[0]));
    v.num_poses = m.pose_ids.size();
    v.pose_ids = m.pose_ids.data();
    v.poses = m.poses.data();
    v.pose_fixed = m.pose_fixed.data();
    v.num_landmarks = m.lm_ids.size();
    v.landmark_ids = m.lm_ids.data();
    v.landmarks = m.lms.data();
    m.sensors.clear();
    m.rig_pose.clear();
    m.rig_lm.clear();
    m.rig_xy.clear();
    m.rig_sensor.clear();
    if (!rig) {
        v.num_observations = m.obs_pose.size();
        v.obs_pose_ids = m.obs_pose.data();
        v.obs_landmark_ids = m.obs_lm.data();
        v.obs_xy = m.obs_xy.data();
    } else {
        for (uint64_t id : used) {
            vba_sensor s{};
            s.camera = to_vba_camera(*find_cam(m, id));
            s.sensor_from_rig[0] = 1.0;
            m.sensors.push_back(s);
        }
        for (size_t i = 0; i < m.obs_pose.size(); ++i) {
            const size_t p = pose_index(m, m.obs_pose[i]);
            // an observation of a pose the problem lacks stays in the list,
            // for vba_optimize to report as MissingPose; its sensor is 0
            const uint64_t cam = p == SIZE_MAX ? used[0] : m.pose_cam[p];
            m.rig_pose.push_back(m.obs_pose[i]);
            m.rig_lm.push_back(m.obs_lm[i]);
            m.rig_xy.push_back(m.obs_xy[2 * i]);
            m.rig_xy.push_back(m.obs_xy[2 * i + 1]);
            m.rig_sensor.push_back((uint32_t)(std::lower_bound(used.begin(), used.end(), cam) - used.begin()));
        }
        v.num_rig_observations = m.rig_pose.size();
        v.rig_pose_ids = m.rig_pose.data();
        v.rig_landmark_ids = m.rig_lm.data();
        v.rig_xy = m.rig_xy.data();
        v.rig_sensor = m.rig_sensor.data();
        v.num_sensors = m.sensors.size();
        v.sensors = m.sensors.data();
    }
    if (m.weighted) v.declined_features |= VBA_FEATURE_OBSERVATION_WEIGHTS;
}

// ba_gpu_replay.rs:43-66.
void keep_local_window(vba_model& m, uint32_t n) {
    const size_t np = m.pose_ids.size();
    std::set<uint64_t> variable(m.pose_ids.begin() + (np - std::min<size_t>(n, np)), m.pose_ids.end());
    std::set<uint64_t> keep_lm;
    for (size_t i = 0; i < m.obs_pose.size(); ++i)
        if (variable.count(m.obs_pose[i])) keep_lm.insert(m.obs_lm[i]);
    vba_model out;
    out.cams = m.cams;
    out.weighted = m.weighted;
    std::set<uint64_t> used;
    for (size_t i = 0; i < m.obs_pose.size(); ++i) {
        if (!keep_lm.count(m.obs_lm[i])) continue;
        out.obs_pose.push_back(m.obs_pose[i]);
        out.obs_lm.push_back(m.obs_lm[i]);
        out.obs_xy.push_back(m.obs_xy[2 * i]);
        out.obs_xy.push_back(m.obs_xy[2 * i + 1]);
        if (m.weighted) out.weights.push_back(m.weights[i]);
        used.insert(m.obs_pose[i]);
    }
    for (size_t i = 0; i < m.lm_ids.size(); ++i)
        if (keep_lm.count(m.lm_ids[i])) {
            out.lm_ids.push_back(m.lm_ids[i]);
            out.lms.insert(out.lms.end(), &m.lms[3 * i], &m.lms[3 * i] + 3);
        }
    for (size_t i = 0; i < np; ++i)
        if (used.count(m.pose_ids[i])) {
            out.pose_ids.push_back(m.pose_ids[i]);
            out.poses.insert(out.poses.end(), &m.poses[7 * i], &m.poses[7 * i] + 7);
            out.pose_fixed.push_back(variable.count(m.pose_ids[i]) ? 0 : 1);
            out.pose_cam.push_back(m.pose_cam[i]);
        }
    // the images' keypoint lists no longer match: written from the
    // observations from now on
    m = std::move(out);
}

const char* model_name(int32_t vba_model_id) {
    for (const auto& m : kModels)
        if (m.vba == vba_model_id) return m.name;
    return nullptr;
}

vba_model* model_from_problem(const vba_problem& pb, const double* weights) {
    if (pb.num_rig_observations)
        fail(VBA_ERR_INVALID_ARGUMENT, "a model holds the monocular part of a problem; it has rig observations");
    const char* name = model_name(pb.camera.model);
    if (!name || pb.camera.num_params > 12) fail(VBA_ERR_UNSUPPORTED_CAMERA_MODEL, "unknown camera model");
    if ((pb.num_poses && (!pb.pose_ids || !pb.poses)) || (pb.num_landmarks && (!pb.landmark_ids || !pb.landmarks)) ||
        (pb.num_observations && (!pb.obs_pose_ids || !pb.obs_landmark_ids || !pb.obs_xy)))
        fail(VBA_ERR_INVALID_ARGUMENT, "null problem array");
    auto* m = new vba_model();
    m->cams.push_back(Cam{1, name, pb.camera.width, pb.camera.height,
                          std::vector<double>(pb.camera.params, pb.camera.params + pb.camera.num_params)});
    // BTreeMap order by id; a repeated id keeps its last entry
    std::map<uint64_t, size_t> pi, li;
    for (size_t i = 0; i < pb.num_poses; ++i) pi[pb.pose_ids[i]] = i;
    for (size_t i = 0; i < pb.num_landmarks; ++i) li[pb.landmark_ids[i]] = i;
    for (const auto& [id, i] : pi) {
        m->pose_ids.push_back(id);
        m->poses.insert(m->poses.end(), pb.poses + 7 * i, pb.poses + 7 * i + 7);
        m->pose_fixed.push_back(pb.pose_fixed && pb.pose_fixed[i] ? 1 : 0);
        m->pose_cam.push_back(1);
    }
    for (const auto& [id, i] : li) {
        m->lm_ids.push_back(id);
        m->lms.insert(m->lms.end(), pb.landmarks + 3 * i, pb.landmarks + 3 * i + 3);
    }
    m->obs_pose.assign(pb.obs_pose_ids, pb.obs_pose_ids + pb.num_observations);
    m->obs_lm.assign(pb.obs_landmark_ids, pb.obs_landmark_ids + pb.num_observations);
    m->obs_xy.assign(pb.obs_xy, pb.obs_xy + 2 * pb.num_observations);
    if (weights) {
        m->weighted = true;
        m->weights.assign(weights, weights + pb.num_observations);
    }
    return m;
}

template <typename F>
int32_t guarded(F&& f) {
    try {
        f();
        return VBA_OK;
    } catch (const IoError& e) {
        g_io_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_io_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_io_error = e.what();
        return VBA_ERR_INVALID_ARGUMENT;
    }
}

}  // namespace

extern "C" {

const char* vba_io_last_error(void) { return g_io_error.c_str(); }

int32_t vba_model_read_ba_problem(const char* path, vba_model** out) {
    return guarded([&] {
        if (!path || !out) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        *out = read_ba_problem(path);
    });
}

int32_t vba_model_write_ba_problem(const vba_model* model, const char* path) {
    return guarded([&] {
        if (!model || !path) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        write_ba_problem(*model, path);
    });
}

int32_t vba_model_read_colmap(const char* dir, int32_t binary, vba_model** out) {
    return guarded([&] {
        if (!dir || !out) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        const std::string d(dir);
        if (binary) {
            auto cams = parse_cameras_bin(read_file(join(d, "cameras.bin")));
            auto pts = parse_points3d_bin(read_file(join(d, "points3D.bin")));
            auto imgs = parse_images_bin(read_file(join(d, "images.bin")));
            *out = model_from_colmap(std::move(cams), std::move(imgs), std::move(pts));
        } else {
            auto cams = parse_cameras_txt(read_file(join(d, "cameras.txt")));
            auto pts = parse_points3d_txt(read_file(join(d, "points3D.txt")));
            auto imgs = parse_images_txt(read_file(join(d, "images.txt")));
            *out = model_from_colmap(std::move(cams), std::move(imgs), std::move(pts));
        }
    });
}

int32_t vba_model_write_colmap(const vba_model* model, const char* dir, int32_t binary) {
    return guarded([&] {
        if (!model || !dir) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        write_colmap(*model, dir, binary != 0);
    });
}

int32_t vba_model_from_problem(const vba_problem* problem, const double* weights, vba_model** out) {
    return guarded([&] {
        if (!problem || !out) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        *out = model_from_problem(*problem, weights);
    });
}

int32_t vba_model_problem(const vba_model* model, vba_problem* out, const double** weights) {
    return guarded([&] {
        if (!model || !out) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        vba_model& m = const_cast<vba_model&>(*model);  // the view buffers only
        build_view(m);
        *out = m.view;
        if (weights) *weights = m.weighted ? m.weights.data() : nullptr;
    });
}

int32_t vba_model_set_state(vba_model* model, const double* poses, const double* landmarks) {
    return guarded([&] {
        if (!model) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        if (poses) std::copy(poses, poses + model->poses.size(), model->poses.begin());
        if (landmarks) std::copy(landmarks, landmarks + model->lms.size(), model->lms.begin());
    });
}

int32_t vba_model_set_fixed(vba_model* model, const uint8_t* fixed) {
    return guarded([&] {
        if (!model || !fixed) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        for (size_t i = 0; i < model->pose_fixed.size(); ++i) model->pose_fixed[i] = fixed[i] ? 1 : 0;
    });
}

int32_t vba_model_keep_local_window(vba_model* model, uint32_t n) {
    return guarded([&] {
        if (!model) fail(VBA_ERR_INVALID_ARGUMENT, "null argument");
        keep_local_window(*model, n);
    });
}

void vba_model_free(vba_model* model) { delete model; }

}  // extern "C"
