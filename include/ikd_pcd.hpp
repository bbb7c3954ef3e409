// ikd_pcd.hpp — header-only reader and writer of PCD point cloud files
// (PCL's format, version 0.7), so the reference's demo flow, loading a .pcd
// map and building the tree from its points
// (examples/ikd_Tree_Search_demo.cpp:64-76, pcl::io::loadPCDFile), works
// without PCL:
//
//   std::vector<ikd_pcl::PointXYZI> cloud;
//   if (ikd_pcd::loadPCDFile("map.pcd", cloud) == -1) { ... }
//   KD_TREE<ikd_pcl::PointXYZI> tree(0.3, 0.6, 0.2);
//   tree.Build(cloud);
//
// Format: a text header (VERSION, FIELDS, SIZE, TYPE, COUNT, WIDTH, HEIGHT,
// VIEWPOINT, POINTS, DATA; '#' lines are comments), then the points as
//   ascii              one point per line, the values in field order;
//   binary             points packed back to back, fields in header order;
//   binary_compressed  uint32 compressed size, uint32 raw size, then an LZF
//                      stream whose raw bytes hold the fields one after
//                      another (all points' first field, then the second, ...).
// Field types: F (float, 4 or 8 bytes), U and I (1, 2, 4 or 8 bytes); a field
// with COUNT > 1 has that many values per point.  The reader fills the point
// type's x, y, z and, where the point type has them, intensity, normal_x,
// normal_y, normal_z and curvature, by field name (every value converted to
// float); other fields are skipped; fields the file lacks stay as the point's
// default constructor left them.  Points are kept as stored (non-finite ones
// included, as PCL keeps them).  The writer stores those same fields.
//
// LZF (Marc Lehmann's liblzf format, which PCL's binary_compressed uses): a
// stream of runs; control byte c < 32: c + 1 literal bytes follow; else a
// back reference of length (c >> 5) + 2 (with c >> 5 == 7 the next byte adds
// to the length), distance ((c & 31) << 8 | next byte) + 1 behind the output.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace ikd_pcd {

// ------------------------------------------------------------------- LZF
// Decodes src into exactly out_len bytes; false on a malformed stream.
inline bool lzf_decompress(const uint8_t* src, size_t src_len, uint8_t* out, size_t out_len) {
    size_t i = 0, o = 0;
    while (i < src_len) {
        const unsigned c = src[i++];
        if (c < 32) {
            const size_t n = c + 1;
            if (i + n > src_len || o + n > out_len) return false;
            std::memcpy(out + o, src + i, n);
            i += n;
            o += n;
            continue;
        }
        size_t len = c >> 5;
        if (len == 7) {
            if (i >= src_len) return false;
            len += src[i++];
        }
        len += 2;
        if (i >= src_len) return false;
        const size_t dist = (((size_t)(c & 31)) << 8 | src[i++]) + 1;
        if (dist > o || o + len > out_len) return false;
        for (size_t k = 0; k < len; ++k, ++o) out[o] = out[o - dist];  // may overlap itself
    }
    return o == out_len;
}

// Greedy encoder: the longest match among the last occurrence of each 3-byte
// prefix (hash table of positions), literals otherwise.  Any valid stream
// decodes the same; this one only has to be valid and reasonably short.
inline std::vector<uint8_t> lzf_compress(const uint8_t* src, size_t n) {
    std::vector<uint8_t> out;
    out.reserve(n + n / 16 + 16);
    std::vector<int64_t> table(1 << 14, -1);
    size_t lit_start = 0;
    auto flush_literals = [&](size_t end) {
        while (lit_start < end) {
            const size_t k = std::min<size_t>(32, end - lit_start);
            out.push_back((uint8_t)(k - 1));
            out.insert(out.end(), src + lit_start, src + lit_start + k);
            lit_start += k;
        }
    };
    size_t i = 0;
    while (i + 2 < n) {
        const uint32_t h = ((uint32_t)src[i] << 16 | (uint32_t)src[i + 1] << 8 | src[i + 2]) * 2654435761u >> 18;
        const int64_t ref = table[h];
        table[h] = (int64_t)i;
        if (ref >= 0 && i - (size_t)ref <= 8192 && std::memcmp(src + ref, src + i, 3) == 0) {
            size_t len = 3;
            const size_t maxlen = std::min<size_t>(264, n - i);
            while (len < maxlen && src[ref + len] == src[i + len]) ++len;
            flush_literals(i);
            const size_t dist = i - (size_t)ref - 1, l = len - 2;
            if (l < 7) {
                out.push_back((uint8_t)((l << 5) | (dist >> 8)));
            } else {
                out.push_back((uint8_t)((7 << 5) | (dist >> 8)));
                out.push_back((uint8_t)(l - 7));
            }
            out.push_back((uint8_t)(dist & 255));
            i += len;
            lit_start = i;
            continue;
        }
        ++i;
    }
    flush_literals(n);
    return out;
}

// ------------------------------------------------------------ point fields
namespace detail {
#define IKD_PCD_HAS(name)                                                                  \
    template <class T, class = void>                                                       \
    struct has_##name : std::false_type {};                                                \
    template <class T>                                                                     \
    struct has_##name<T, std::void_t<decltype(std::declval<T&>().name)>> : std::true_type {};
IKD_PCD_HAS(intensity)
IKD_PCD_HAS(normal_x)
IKD_PCD_HAS(normal_y)
IKD_PCD_HAS(normal_z)
IKD_PCD_HAS(curvature)
#undef IKD_PCD_HAS

// Field index the reader knows: 0-2 x y z, 3 intensity, 4-6 normal xyz, 7 curvature, -1 other.
inline int known_field(const std::string& f) {
    static const char* names[] = {"x", "y", "z", "intensity", "normal_x", "normal_y", "normal_z", "curvature"};
    for (int i = 0; i < 8; ++i)
        if (f == names[i]) return i;
    return -1;
}

template <class P>
void set_field(P& p, int which, float v) {
    switch (which) {
        case 0: p.x = v; break;
        case 1: p.y = v; break;
        case 2: p.z = v; break;
        case 3: if constexpr (has_intensity<P>::value) p.intensity = v; break;
        case 4: if constexpr (has_normal_x<P>::value) p.normal_x = v; break;
        case 5: if constexpr (has_normal_y<P>::value) p.normal_y = v; break;
        case 6: if constexpr (has_normal_z<P>::value) p.normal_z = v; break;
        case 7: if constexpr (has_curvature<P>::value) p.curvature = v; break;
        default: break;
    }
}

template <class P>
std::vector<std::pair<std::string, float P::*>> fields_of() {
    std::vector<std::pair<std::string, float P::*>> f = {{"x", &P::x}, {"y", &P::y}, {"z", &P::z}};
    if constexpr (has_intensity<P>::value) f.push_back({"intensity", &P::intensity});
    if constexpr (has_normal_x<P>::value && has_normal_y<P>::value && has_normal_z<P>::value) {
        f.push_back({"normal_x", &P::normal_x});
        f.push_back({"normal_y", &P::normal_y});
        f.push_back({"normal_z", &P::normal_z});
    }
    if constexpr (has_curvature<P>::value) f.push_back({"curvature", &P::curvature});
    return f;
}

inline bool read_value(const uint8_t* b, char type, int size, double* v) {
    switch (type) {
        case 'F':
            if (size == 4) { float f; std::memcpy(&f, b, 4); *v = f; return true; }
            if (size == 8) { double d; std::memcpy(&d, b, 8); *v = d; return true; }
            return false;
        case 'U':
            if (size == 1) { *v = b[0]; return true; }
            if (size == 2) { uint16_t u; std::memcpy(&u, b, 2); *v = u; return true; }
            if (size == 4) { uint32_t u; std::memcpy(&u, b, 4); *v = u; return true; }
            if (size == 8) { uint64_t u; std::memcpy(&u, b, 8); *v = (double)u; return true; }
            return false;
        case 'I':
            if (size == 1) { *v = (int8_t)b[0]; return true; }
            if (size == 2) { int16_t u; std::memcpy(&u, b, 2); *v = u; return true; }
            if (size == 4) { int32_t u; std::memcpy(&u, b, 4); *v = u; return true; }
            if (size == 8) { int64_t u; std::memcpy(&u, b, 8); *v = (double)u; return true; }
            return false;
        default:
            return false;
    }
}
}  // namespace detail

struct Header {
    std::vector<std::string> fields;
    std::vector<int> size, count;
    std::vector<char> type;
    long long width = 0, height = 1, points = -1;
    std::string data;  // ascii | binary | binary_compressed
    size_t data_offset = 0;
};

// Parses the header of the file's bytes; false (with *err) when malformed.
inline bool parse_header(const std::string& bytes, Header* h, std::string* err) {
    size_t pos = 0;
    auto fail = [&](const std::string& m) {
        if (err) *err = m;
        return false;
    };
    while (pos < bytes.size()) {
        size_t eol = bytes.find('\n', pos);
        if (eol == std::string::npos) return fail("PCD header ends without DATA");
        std::string line = bytes.substr(pos, eol - pos);
        pos = eol + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "VERSION" || key == "VIEWPOINT") continue;
        if (key == "FIELDS" || key == "COLUMNS") {
            std::string f;
            while (ss >> f) h->fields.push_back(f);
        } else if (key == "SIZE") {
            int v;
            while (ss >> v) h->size.push_back(v);
        } else if (key == "TYPE") {
            std::string v;
            while (ss >> v) h->type.push_back(v.empty() ? '?' : v[0]);
        } else if (key == "COUNT") {
            int v;
            while (ss >> v) h->count.push_back(v);
        } else if (key == "WIDTH") {
            ss >> h->width;
        } else if (key == "HEIGHT") {
            ss >> h->height;
        } else if (key == "POINTS") {
            ss >> h->points;
        } else if (key == "DATA") {
            ss >> h->data;
            h->data_offset = pos;
            break;
        } else {
            return fail("unknown PCD header line: " + line);
        }
    }
    const size_t nf = h->fields.size();
    if (h->data.empty()) return fail("PCD header has no DATA line");
    if (nf == 0 || h->size.size() != nf || h->type.size() != nf) return fail("PCD FIELDS/SIZE/TYPE disagree");
    if (h->count.empty()) h->count.assign(nf, 1);
    if (h->count.size() != nf) return fail("PCD COUNT disagrees with FIELDS");
    if (h->points < 0) h->points = h->width * h->height;
    if (h->points != h->width * h->height) return fail("PCD POINTS != WIDTH * HEIGHT");
    for (size_t i = 0; i < nf; ++i) {
        double v;
        uint8_t z[8] = {};
        if (h->count[i] < 1 || !detail::read_value(z, h->type[i], h->size[i], &v))
            return fail("unsupported PCD field type/size for " + h->fields[i]);
    }
    return true;
}

// Reads every point of the file into out (cleared first).  Returns 0, or -1
// with *err set.
template <class PointT, class Alloc>
int load(const std::string& path, std::vector<PointT, Alloc>& out, std::string* err = nullptr) {
    out.clear();
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        if (err) *err = "cannot open " + path;
        return -1;
    }
    std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Header h;
    if (!parse_header(bytes, &h, err)) return -1;
    const size_t nf = h.fields.size(), n = (size_t)h.points;
    std::vector<int> known(nf);
    std::vector<size_t> off(nf);  // byte offset of each field within a packed point
    size_t psize = 0;
    for (size_t i = 0; i < nf; ++i) {
        known[i] = detail::known_field(h.fields[i]);
        off[i] = psize;
        psize += (size_t)h.size[i] * h.count[i];
    }
    out.resize(n);
    auto fail = [&](const std::string& m) {
        if (err) *err = m;
        out.clear();
        return -1;
    };
    const uint8_t* body = reinterpret_cast<const uint8_t*>(bytes.data()) + h.data_offset;
    const size_t body_len = bytes.size() - h.data_offset;
    if (h.data == "ascii") {
        std::istringstream ss(bytes.substr(h.data_offset));
        for (size_t p = 0; p < n; ++p) {
            for (size_t i = 0; i < nf; ++i) {
                for (int c = 0; c < h.count[i]; ++c) {
                    std::string tok;
                    if (!(ss >> tok)) return fail("PCD ascii data ends early");
                    // through double, as the binary forms convert: one rounding to float
                    const float v = (tok == "nan" || tok == "NaN") ? NAN : (float)std::strtod(tok.c_str(), nullptr);
                    if (c == 0) detail::set_field(out[p], known[i], v);
                }
            }
        }
        return 0;
    }
    if (h.data == "binary") {
        if (body_len < n * psize) return fail("PCD binary data ends early");
        for (size_t p = 0; p < n; ++p) {
            const uint8_t* b = body + p * psize;
            for (size_t i = 0; i < nf; ++i) {
                double v;
                if (known[i] >= 0 && detail::read_value(b + off[i], h.type[i], h.size[i], &v))
                    detail::set_field(out[p], known[i], (float)v);
            }
        }
        return 0;
    }
    if (h.data == "binary_compressed") {
        if (body_len < 8) return fail("PCD compressed data ends early");
        uint32_t csize, rsize;
        std::memcpy(&csize, body, 4);
        std::memcpy(&rsize, body + 4, 4);
        if (body_len < 8 + (size_t)csize) return fail("PCD compressed data ends early");
        if ((size_t)rsize != n * psize) return fail("PCD compressed size disagrees with the fields");
        std::vector<uint8_t> raw(rsize);
        if (rsize > 0 && !lzf_decompress(body + 8, csize, raw.data(), raw.size()))
            return fail("PCD LZF stream is malformed");
        // field-major: field i's values for all points start at n * off[i]
        for (size_t i = 0; i < nf; ++i) {
            if (known[i] < 0) continue;
            const size_t stride = (size_t)h.size[i] * h.count[i];
            const uint8_t* col = raw.data() + n * off[i];
            for (size_t p = 0; p < n; ++p) {
                double v;
                if (detail::read_value(col + p * stride, h.type[i], h.size[i], &v))
                    detail::set_field(out[p], known[i], (float)v);
            }
        }
        return 0;
    }
    return fail("unknown PCD DATA kind: " + h.data);
}

// pcl::io::loadPCDFile's shape: 0 on success, -1 on failure (message on stderr).
template <class PointT, class Alloc>
int loadPCDFile(const std::string& path, std::vector<PointT, Alloc>& out) {
    std::string err;
    const int rc = load(path, out, &err);
    if (rc != 0) std::fprintf(stderr, "ikd_pcd: %s\n", err.c_str());
    return rc;
}
// ... and for cloud types holding a `points` vector (pcl::PointCloud's shape).
template <class Cloud>
auto loadPCDFile(const std::string& path, Cloud& cloud) -> decltype(cloud.points.clear(), int()) {
    const int rc = loadPCDFile(path, cloud.points);
    if (rc == 0) {
        cloud.width = (decltype(cloud.width))cloud.points.size();
        cloud.height = 1;
    }
    return rc;
}

// Writes the point type's known fields (x y z, intensity, normals,
// curvature) as float32.  mode: "ascii", "binary" or "binary_compressed".
template <class PointT, class Alloc>
int save(const std::string& path, const std::vector<PointT, Alloc>& pts, const std::string& mode = "binary",
         std::string* err = nullptr) {
    const auto f = detail::fields_of<PointT>();
    const size_t nf = f.size(), n = pts.size();
    std::ostringstream hdr;
    hdr << "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS";
    for (auto& e : f) hdr << ' ' << e.first;
    hdr << "\nSIZE";
    for (size_t i = 0; i < nf; ++i) hdr << " 4";
    hdr << "\nTYPE";
    for (size_t i = 0; i < nf; ++i) hdr << " F";
    hdr << "\nCOUNT";
    for (size_t i = 0; i < nf; ++i) hdr << " 1";
    hdr << "\nWIDTH " << n << "\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS " << n << "\nDATA " << mode << '\n';
    std::ofstream o(path, std::ios::binary);
    if (!o) {
        if (err) *err = "cannot write " + path;
        return -1;
    }
    o << hdr.str();
    if (mode == "ascii") {
        char buf[64];
        for (const PointT& p : pts) {
            for (size_t i = 0; i < nf; ++i) {
                std::snprintf(buf, sizeof buf, "%.9g", (double)(p.*(f[i].second)));
                o << (i ? " " : "") << buf;
            }
            o << '\n';
        }
    } else if (mode == "binary") {
        std::vector<float> row(nf);
        for (const PointT& p : pts) {
            for (size_t i = 0; i < nf; ++i) row[i] = p.*(f[i].second);
            o.write(reinterpret_cast<const char*>(row.data()), nf * 4);
        }
    } else if (mode == "binary_compressed") {
        std::vector<float> raw(n * nf);
        for (size_t i = 0; i < nf; ++i)
            for (size_t p = 0; p < n; ++p) raw[i * n + p] = pts[p].*(f[i].second);
        const auto c = lzf_compress(reinterpret_cast<const uint8_t*>(raw.data()), raw.size() * 4);
        const uint32_t cs = (uint32_t)c.size(), rs = (uint32_t)(raw.size() * 4);
        o.write(reinterpret_cast<const char*>(&cs), 4);
        o.write(reinterpret_cast<const char*>(&rs), 4);
        o.write(reinterpret_cast<const char*>(c.data()), c.size());
    } else {
        if (err) *err = "unknown PCD DATA kind: " + mode;
        return -1;
    }
    return o ? 0 : -1;
}

}  // namespace ikd_pcd
