// polyMesh reader (host code): the files the reference reads in
// src/mesh_geometry_and_topology.f90:409-605 -- points, faces, owner,
// neighbour and the simplified boundary file -- in either of the two layouts
// it accepts:
//   * OpenFOAM ASCII: FoamFile header, then `N ( ... )` lists, 0-based labels;
//     nCells from the owner header's note (as the reference, :516-527) or,
//     without a note, from the largest owner label;
//   * the reference's native layout: owner starts with numNodes, numCells,
//     numFaces, numInnerFaces, labels are 1-based, one face per line as
//     `count n1 n2 ...` (:473-480).
// The label and coordinate lists are parsed by several threads: the body of a
// list is cut at entry boundaries into one chunk per thread, each thread
// parses its chunk into its own buffer, and the buffers are joined in order,
// so the result does not depend on the thread count.  Faces may have any
// number of vertices (the reference's OpenFOAM path reads single-digit counts
// only, :291,297).
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fcap.h"

namespace {

struct ReadError {
    fc_status code;
    std::string msg;
};

[[noreturn]] void fail(fc_status code, const std::string &m) { throw ReadError{code, m}; }

std::string slurp(const std::string &path) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fail(FC_ERR_ARGUMENT, "cannot open " + path);
    std::string s;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long len = std::ftell(f);
        if (len > 0) {
            s.resize((size_t)len);
            std::rewind(f);
            const size_t got = std::fread(&s[0], 1, (size_t)len, f);
            s.resize(got);
        }
    }
    std::fclose(f);
    return s;
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\f' || c == '\v'; }

// position after the FoamFile header: skips /* */ and // comments and the
// FoamFile { ... } dictionary; returns 0 when the text has no FoamFile header
size_t skip_foam_header(const std::string &t, std::string *note) {
    size_t i = 0, n = t.size();
    bool seen = false;
    while (i < n) {
        if (is_space(t[i])) { ++i; continue; }
        if (t.compare(i, 2, "/*") == 0) {
            size_t e = t.find("*/", i + 2);
            if (e == std::string::npos) fail(FC_ERR_ARGUMENT, "unterminated comment");
            i = e + 2;
            continue;
        }
        if (t.compare(i, 2, "//") == 0) {
            size_t e = t.find('\n', i);
            i = e == std::string::npos ? n : e + 1;
            continue;
        }
        if (t.compare(i, 8, "FoamFile") == 0) {
            size_t b = t.find('{', i), e = t.find('}', i);
            if (b == std::string::npos || e == std::string::npos || e < b) fail(FC_ERR_ARGUMENT, "bad FoamFile header");
            const std::string dict = t.substr(b + 1, e - b - 1);
            if (dict.find("binary") != std::string::npos && dict.find("ascii") == std::string::npos)
                fail(FC_ERR_ARGUMENT, "binary polyMesh files are not supported (write them as ascii)");
            size_t q = dict.find("note");
            if (note && q != std::string::npos) {
                size_t a = dict.find('"', q), z = a == std::string::npos ? a : dict.find('"', a + 1);
                if (a != std::string::npos && z != std::string::npos) *note = dict.substr(a + 1, z - a - 1);
            }
            i = e + 1;
            seen = true;
            continue;
        }
        break;
    }
    return seen ? i : 0;
}

long long parse_ll(const char *&p, const char *end, const char *what) {
    while (p < end && is_space(*p)) ++p;
    long long v = 0;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) fail(FC_ERR_ARGUMENT, std::string("expected an integer in ") + what);
    p = r.ptr;
    return v;
}

// `N (` of a list; returns N and leaves p after the '('
long long list_head(const char *&p, const char *end, const char *what) {
    long long cnt = parse_ll(p, end, what);
    while (p < end && is_space(*p)) ++p;
    if (p >= end || *p != '(') fail(FC_ERR_ARGUMENT, std::string("expected '(' after the count in ") + what);
    ++p;
    return cnt;
}

// Cut [b, e) into up to `parts` chunks whose boundaries lie just after a
// character for which `cut` is true (entry boundaries).
template <class Cut>
std::vector<const char *> chunks(const char *b, const char *e, int parts, Cut cut) {
    std::vector<const char *> pos{b};
    const size_t len = (size_t)(e - b);
    for (int k = 1; k < parts; ++k) {
        const char *q = b + len * k / parts;
        if (q <= pos.back()) continue;
        while (q < e && !cut(q[-1])) ++q;
        if (q < e && q > pos.back()) pos.push_back(q);
    }
    pos.push_back(e);
    return pos;
}

template <class F>
void run_parallel(int nthreads, int parts, F &&f) {
    if (nthreads <= 1 || parts <= 1) {
        for (int k = 0; k < parts; ++k) f(k);
        return;
    }
    std::vector<std::thread> th;
    std::vector<ReadError> errs(parts);
    std::vector<char> bad(parts, 0);
    for (int k = 0; k < parts; ++k)
        th.emplace_back([&, k]() {
            try {
                f(k);
            } catch (const ReadError &e) {
                errs[k] = e;
                bad[k] = 1;
            }
        });
    for (auto &t : th) t.join();
    for (int k = 0; k < parts; ++k)
        if (bad[k]) throw errs[k];
}

// flat numbers of [b, e) with '(' and ')' as separators
template <class T>
void parse_numbers(const char *b, const char *e, std::vector<T> &out, const char *what) {
    const char *p = b;
    while (true) {
        while (p < e && (is_space(*p) || *p == '(' || *p == ')')) ++p;
        if (p >= e) break;
        T v{};
        auto r = std::from_chars(p, e, v);
        if (r.ec != std::errc()) fail(FC_ERR_ARGUMENT, std::string("bad number in ") + what);
        out.push_back(v);
        p = r.ptr;
    }
}

template <class T>
std::vector<T> parse_numbers_mt(const char *b, const char *e, int nthreads, const char *what, bool paren_cut) {
    const int parts = std::max(1, nthreads) * 4;
    auto pos = chunks(b, e, (e - b) < (1 << 16) ? 1 : parts,
                      [&](char c) { return paren_cut ? c == ')' : is_space(c); });
    const int np = (int)pos.size() - 1;
    std::vector<std::vector<T>> part(np);
    run_parallel(nthreads, np, [&](int k) { parse_numbers<T>(pos[k], pos[k + 1], part[k], what); });
    size_t tot = 0;
    for (auto &v : part) tot += v.size();
    std::vector<T> out;
    out.reserve(tot);
    for (auto &v : part) out.insert(out.end(), v.begin(), v.end());
    return out;
}

// end of a list body: the last ')' of the text
const char *list_end(const std::string &t, const char *p, const char *what) {
    size_t e = t.rfind(')');
    if (e == std::string::npos || t.data() + e < p) fail(FC_ERR_ARGUMENT, std::string("unterminated list in ") + what);
    return t.data() + e;
}

// faces: entries `cnt ( n1 ... ncnt )` (foam) or lines `cnt n1 ... ncnt` (native)
void parse_faces(const char *b, const char *e, bool foam, int nthreads, std::vector<int> &ptr,
                 std::vector<int> &nodes, int base) {
    const int parts = std::max(1, nthreads) * 4;
    auto pos = chunks(b, e, (e - b) < (1 << 16) ? 1 : parts, [&](char c) { return foam ? c == ')' : c == '\n'; });
    const int np = (int)pos.size() - 1;
    std::vector<std::vector<int>> cnts(np), nds(np);
    run_parallel(nthreads, np, [&](int k) {
        const char *p = pos[k], *q = pos[k + 1];
        while (true) {
            while (p < q && is_space(*p)) ++p;
            if (p >= q) break;
            int c = 0;
            auto r = std::from_chars(p, q, c);
            if (r.ec != std::errc() || c < 3) fail(FC_ERR_ARGUMENT, "bad vertex count in faces");
            p = r.ptr;
            if (foam) {
                while (p < q && is_space(*p)) ++p;
                if (p >= q || *p != '(') fail(FC_ERR_ARGUMENT, "expected '(' in faces");
                ++p;
            }
            for (int j = 0; j < c; ++j) {
                while (p < q && is_space(*p)) ++p;
                int v = 0;
                auto r2 = std::from_chars(p, q, v);
                if (r2.ec != std::errc()) fail(FC_ERR_ARGUMENT, "bad vertex label in faces");
                nds[k].push_back(v - base);
                p = r2.ptr;
            }
            if (foam) {
                while (p < q && is_space(*p)) ++p;
                if (p >= q || *p != ')') fail(FC_ERR_ARGUMENT, "expected ')' in faces");
                ++p;
            }
            cnts[k].push_back(c);
        }
    });
    ptr.assign(1, 0);
    size_t tot = 0, totf = 0;
    for (int k = 0; k < np; ++k) {
        tot += nds[k].size();
        totf += cnts[k].size();
    }
    ptr.reserve(totf + 1);
    nodes.clear();
    nodes.reserve(tot);
    for (int k = 0; k < np; ++k) {
        for (int c : cnts[k]) ptr.push_back(ptr.back() + c);
        nodes.insert(nodes.end(), nds[k].begin(), nds[k].end());
    }
}

bool known_bctype(const std::string &t) {
    static const char *kTypes[] = {"inlet",     "outlet",    "symmetry", "wall",   "wallIsoth",
                                   "wallAdiab", "wallQFlux", "prOutlet", "domain", "cyclic"};
    for (const char *k : kTypes)
        if (t == k) return true;
    return false;
}

}  // namespace

struct fc_polymesh {
    int num_nodes = 0, num_faces = 0, num_inner = 0, num_cells = 0;
    std::vector<double> xyz;
    std::vector<int> face_ptr, face_nodes, owner, neighbour;   // 0-based labels
    struct Patch {
        std::string type;
        int nfaces, start;
    };
    std::vector<Patch> patches;
};

namespace {

void read_all(const std::string &dir, int nthreads, fc_polymesh &m) {
    // the four files are read concurrently, then each is parsed by all threads
    std::string files[4];
    static const char *kNames[4] = {"owner", "neighbour", "points", "faces"};
    run_parallel(std::min(nthreads, 4), 4, [&](int k) { files[k] = slurp(dir + "/" + kNames[k]); });
    const std::string &owner_t = files[0];
    std::string note;
    const size_t ob = skip_foam_header(owner_t, &note);
    const bool foam = ob != 0;
    int ncells_decl = -1;
    // owner
    {
        const char *p = owner_t.data() + ob, *end = owner_t.data() + owner_t.size();
        if (foam) {
            const long long nf = list_head(p, end, "owner");
            const char *e = list_end(owner_t, p, "owner");
            m.owner = parse_numbers_mt<int>(p, e, nthreads, "owner", false);
            if ((long long)m.owner.size() != nf) fail(FC_ERR_ARGUMENT, "owner: count differs from its header");
            m.num_faces = (int)nf;
            size_t q = note.find("nCells");
            if (q != std::string::npos) {
                const char *a = note.data() + q + 6, *z = note.data() + note.size();
                while (a < z && (*a == ':' || is_space(*a))) ++a;
                int v = 0;
                if (std::from_chars(a, z, v).ec == std::errc()) ncells_decl = v;
            }
        } else {
            m.num_nodes = (int)parse_ll(p, end, "owner");
            ncells_decl = (int)parse_ll(p, end, "owner");
            m.num_faces = (int)parse_ll(p, end, "owner");
            m.num_inner = (int)parse_ll(p, end, "owner");
            m.owner = parse_numbers_mt<int>(p, end, nthreads, "owner", false);
            if ((int)m.owner.size() != m.num_faces) fail(FC_ERR_ARGUMENT, "owner: count differs from numFaces");
            for (int &v : m.owner) v -= 1;
        }
    }
    // neighbour
    {
        const std::string &t = files[1];
        const size_t b = skip_foam_header(t, nullptr);
        const char *p = t.data() + b, *end = t.data() + t.size();
        if (foam) {
            const long long ni = list_head(p, end, "neighbour");
            const char *e = list_end(t, p, "neighbour");
            m.neighbour = parse_numbers_mt<int>(p, e, nthreads, "neighbour", false);
            if ((long long)m.neighbour.size() != ni) fail(FC_ERR_ARGUMENT, "neighbour: count differs from its header");
            m.num_inner = (int)ni;
        } else {
            m.neighbour = parse_numbers_mt<int>(p, end, nthreads, "neighbour", false);
            if ((int)m.neighbour.size() != m.num_inner) fail(FC_ERR_ARGUMENT, "neighbour: count differs from numInnerFaces");
            for (int &v : m.neighbour) v -= 1;
        }
    }
    // points
    {
        const std::string &t = files[2];
        const size_t b = skip_foam_header(t, nullptr);
        const char *p = t.data() + b, *end = t.data() + t.size();
        if (foam) {
            const long long nn = list_head(p, end, "points");
            const char *e = list_end(t, p, "points");
            m.xyz = parse_numbers_mt<double>(p, e, nthreads, "points", true);
            if ((long long)m.xyz.size() != 3 * nn) fail(FC_ERR_ARGUMENT, "points: count differs from its header");
            m.num_nodes = (int)nn;
        } else {
            m.xyz = parse_numbers_mt<double>(p, end, nthreads, "points", false);
            if ((long long)m.xyz.size() != 3LL * m.num_nodes) fail(FC_ERR_ARGUMENT, "points: count differs from numNodes");
        }
    }
    // faces
    {
        const std::string &t = files[3];
        const size_t b = skip_foam_header(t, nullptr);
        const char *p = t.data() + b, *end = t.data() + t.size();
        if (foam) {
            const long long nf = list_head(p, end, "faces");
            const char *e = list_end(t, p, "faces");
            parse_faces(p, e, true, nthreads, m.face_ptr, m.face_nodes, 0);
            if ((long long)m.face_ptr.size() - 1 != nf) fail(FC_ERR_ARGUMENT, "faces: count differs from its header");
        } else {
            parse_faces(p, end, false, nthreads, m.face_ptr, m.face_nodes, 1);
        }
        if ((int)m.face_ptr.size() - 1 != m.num_faces) fail(FC_ERR_ARGUMENT, "faces: count differs from owner");
    }
    // boundary (the reference's simplified file: '#' comment, `bctype nFaces startFace` lines)
    {
        std::ifstream f(dir + "/boundary");
        if (!f) fail(FC_ERR_ARGUMENT, "cannot open " + dir + "/boundary");
        std::string line;
        while (std::getline(f, line)) {
            size_t a = line.find_first_not_of(" \t\r");
            if (a == std::string::npos || line[a] == '#') continue;
            std::istringstream ls(line);
            fc_polymesh::Patch pt;
            if (!(ls >> pt.type >> pt.nfaces >> pt.start))
                fail(FC_ERR_ARGUMENT, "boundary: expected 'bctype nFaces startFace' lines, got: " + line);
            if (!known_bctype(pt.type)) fail(FC_ERR_ARGUMENT, "boundary: unknown bctype " + pt.type);
            m.patches.push_back(pt);
        }
    }
    // checks
    const int nf = m.num_faces, nif = m.num_inner, nn = m.num_nodes;
    if (nif < 0 || nif > nf) fail(FC_ERR_STRUCTURE, "more inner faces than faces");
    int maxo = -1;
    for (int i = 0; i < nf; ++i) {
        if (m.owner[i] < 0) fail(FC_ERR_STRUCTURE, "negative owner label");
        maxo = std::max(maxo, m.owner[i]);
    }
    m.num_cells = ncells_decl > 0 ? ncells_decl : maxo + 1;
    if (maxo >= m.num_cells) fail(FC_ERR_STRUCTURE, "owner label beyond nCells");
    for (int i = 0; i < nif; ++i)
        if (m.neighbour[i] < 0 || m.neighbour[i] >= m.num_cells) fail(FC_ERR_STRUCTURE, "neighbour label out of range");
    for (int v : m.face_nodes)
        if (v < 0 || v >= nn) fail(FC_ERR_STRUCTURE, "face vertex label out of range");
    long long covered = 0;
    for (auto &pt : m.patches) {
        if (pt.nfaces < 0 || pt.start < nif || (long long)pt.start + pt.nfaces > nf)
            fail(FC_ERR_STRUCTURE, "boundary: patch " + pt.type + " lies outside the boundary faces");
        covered += pt.nfaces;
    }
    if (covered != nf - nif) fail(FC_ERR_STRUCTURE, "boundary: patches do not cover the boundary faces");
}

int threads_or_default(int n) {
    if (n > 0) return n;
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

}  // namespace

extern "C" fc_status fc_polymesh_read(const char *dir, int nthreads, fc_polymesh **out, char *err, int errlen) {
    if (err && errlen > 0) err[0] = 0;
    if (!dir || !out) return FC_ERR_ARGUMENT;
    *out = nullptr;
    fc_polymesh *m = new fc_polymesh();
    try {
        read_all(dir, threads_or_default(nthreads), *m);
    } catch (const ReadError &e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.msg.c_str());
        delete m;
        return e.code;
    } catch (const std::exception &e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.what());
        delete m;
        return FC_ERR_ALLOC;
    }
    *out = m;
    return FC_OK;
}

extern "C" fc_status fc_polymesh_sizes(const fc_polymesh *m, int *numNodes, int *numFaces, int *numInnerFaces,
                                       int *numCells, int *numFaceNodes, int *numPatches) {
    if (!m) return FC_ERR_ARGUMENT;
    if (numNodes) *numNodes = m->num_nodes;
    if (numFaces) *numFaces = m->num_faces;
    if (numInnerFaces) *numInnerFaces = m->num_inner;
    if (numCells) *numCells = m->num_cells;
    if (numFaceNodes) *numFaceNodes = (int)m->face_nodes.size();
    if (numPatches) *numPatches = (int)m->patches.size();
    return FC_OK;
}

extern "C" fc_status fc_polymesh_arrays(const fc_polymesh *m, double *xyz, int *face_ptr, int *face_nodes,
                                        int *owner, int *neighbour) {
    if (!m) return FC_ERR_ARGUMENT;
    if (xyz) std::memcpy(xyz, m->xyz.data(), sizeof(double) * m->xyz.size());
    if (face_ptr) std::memcpy(face_ptr, m->face_ptr.data(), sizeof(int) * m->face_ptr.size());
    if (face_nodes)
        for (size_t i = 0; i < m->face_nodes.size(); ++i) face_nodes[i] = m->face_nodes[i] + 1;
    if (owner)
        for (size_t i = 0; i < m->owner.size(); ++i) owner[i] = m->owner[i] + 1;
    if (neighbour)
        for (size_t i = 0; i < m->neighbour.size(); ++i) neighbour[i] = m->neighbour[i] + 1;
    return FC_OK;
}

extern "C" fc_status fc_polymesh_patch(const fc_polymesh *m, int index, char *bctype, int len, int *nFaces,
                                       int *startFace) {
    if (!m || index < 0 || index >= (int)m->patches.size()) return FC_ERR_ARGUMENT;
    const auto &p = m->patches[index];
    if (bctype && len > 0) std::snprintf(bctype, (size_t)len, "%s", p.type.c_str());
    if (nFaces) *nFaces = p.nfaces;
    if (startFace) *startFace = p.start;
    return FC_OK;
}

extern "C" void fc_polymesh_free(fc_polymesh *m) { delete m; }
