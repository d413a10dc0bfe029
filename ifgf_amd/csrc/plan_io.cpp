// plan_io.cpp -- the plan of an operator written to and read from a file, so that a later
// process skips the octree build, the interaction lists, the cone marking and the upward
// tables (octree.hpp:202-816 and ifgfoperator.hpp:65-106 redo all of them in every init).
//
// Layout (native byte order, checked by a marker): an 8-byte magic, the format version, the
// byte-order marker, then the payload: the operator parameters the plan was made for, every
// array of the Plan as (count, raw elements), the launch shape in use; last, the payload's
// byte count and a 64-bit checksum (FNV-1a steps over 8-byte words).  The reader checks the
// magic, version, marker, length and checksum, then every count against the bytes left,
// then the sizes and indices of the arrays, before it hands out anything.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "plan.hpp"
#include "plan_io.hpp"

namespace ifgf
{
namespace
{
constexpr char MAGIC[8] = {'I', 'F', 'G', 'F', 'P', 'L', 'A', 'N'};
constexpr uint32_t VERSION = 1;
constexpr uint32_t ORDER_MARK = 0x01020304u;

// checksum of a byte stream fed in pieces of any size: FNV-1a steps over its 8-byte words
// (native order), the last partial word zero-padded, then the length
struct Checksum
{
    uint64_t h = 0xcbf29ce484222325ull, bytes = 0, carry = 0;
    int nc = 0; // bytes in carry

    void mix(uint64_t w) { h = (h ^ w) * 0x100000001b3ull; }
    void add(const unsigned char* p, size_t n)
    {
        bytes += n;
        while (n && nc)
        {
            carry |= uint64_t(*p++) << (8 * nc);
            n--;
            if (++nc == 8)
            {
                mix(carry);
                carry = 0;
                nc = 0;
            }
        }
        for (; n >= 8; n -= 8, p += 8)
        {
            uint64_t w;
            std::memcpy(&w, p, 8);
            mix(w);
        }
        // here n < 8 and (nc == 0 or n == 0)
        for (size_t i = 0; i < n; i++)
            carry |= uint64_t(p[i]) << (8 * ((size_t(nc) + i) & 7));
        nc += int(n);
    }
    uint64_t done()
    {
        if (nc)
            mix(carry);
        mix(bytes);
        return h;
    }
};

struct Writer
{
    FILE* f = nullptr;
    Checksum sum;
    bool ok = true;

    void raw(const void* p, size_t n)
    {
        sum.add(static_cast<const unsigned char*>(p), n);
        if (n && std::fwrite(p, 1, n, f) != n)
            ok = false;
    }
    template <typename T> void put(const T& v)
    {
        static_assert(std::is_trivially_copyable<T>::value, "raw value");
        raw(&v, sizeof(T));
    }
    template <typename T> void vec(const std::vector<T>& v)
    {
        static_assert(std::is_trivially_copyable<T>::value, "raw elements");
        put(int64_t(v.size()));
        raw(v.data(), v.size() * sizeof(T));
    }
};

struct Reader
{
    const unsigned char* p = nullptr;
    size_t left = 0;
    bool ok = true;

    void raw(void* out, size_t n)
    {
        if (!ok || n > left)
        {
            ok = false;
            return;
        }
        std::memcpy(out, p, n);
        p += n;
        left -= n;
    }
    template <typename T> void get(T& v) { raw(&v, sizeof(T)); }
    template <typename T> void vec(std::vector<T>& v)
    {
        int64_t n = -1;
        get(n);
        if (!ok || n < 0 || uint64_t(n) > left / sizeof(T))
        {
            ok = false;
            return;
        }
        v.resize(size_t(n));
        raw(v.data(), size_t(n) * sizeof(T));
    }
};

void walk_tree(Writer& w, const Tree& t)
{
    w.put(t.n);
    w.vec(t.pts);
    w.vec(t.perm);
    w.put(t.lo);
    w.put(t.hi);
    w.put(t.levels);
    w.vec(t.level_off);
    w.vec(t.begin);
    w.vec(t.end);
    w.vec(t.parent);
    w.vec(t.child);
    w.vec(t.leaf);
    w.vec(t.blo);
    w.vec(t.bhi);
    w.vec(t.ctr);
    w.vec(t.side);
}

void walk_tree(Reader& r, Tree& t)
{
    r.get(t.n);
    r.vec(t.pts);
    r.vec(t.perm);
    r.get(t.lo);
    r.get(t.hi);
    r.get(t.levels);
    r.vec(t.level_off);
    r.vec(t.begin);
    r.vec(t.end);
    r.vec(t.parent);
    r.vec(t.child);
    r.vec(t.leaf);
    r.vec(t.blo);
    r.vec(t.bhi);
    r.vec(t.ctr);
    r.vec(t.side);
}

void walk_up(Writer& w, const UpLevel& u)
{
    w.put(uint8_t(u.valid));
    w.vec(u.cellmap);
    w.vec(u.entries);
}

void walk_up(Reader& r, UpLevel& u)
{
    uint8_t v = 0;
    r.get(v);
    u.valid = v != 0;
    r.vec(u.cellmap);
    r.vec(u.entries);
}

// the plan's arrays in one fixed order (both directions use this list)
#define IFGF_PLAN_ARRAYS(X)                                                                   \
    X(normals) X(far) X(near) X(has_far_anc) X(parent_has_far) X(dom_lo) X(dom_hi) X(nc)      \
        X(nf) X(hc) X(hf) X(table_off) X(cone_off) X(cone_box) X(cone_cell) X(cone_mask)      \
            X(cone_fine) X(table) X(fine_off) X(fine_ids) X(box_up_part) X(box_k)             \
                X(part_src_off) X(part_tgt_off)

void walk_plan(Writer& w, const Plan& P)
{
    w.put(P.K.kind);
    w.put(P.K.k.real());
    w.put(P.K.k.imag());
    w.put(P.K.kappa);
    w.put(P.K.dx);
    w.put(P.K.tol);
    w.put(P.base_els);
    w.put(P.base_order);
    w.put(P.ho_order);
    w.put(P.lo_order);
    w.put(P.factor);
    walk_tree(w, P.S);
    walk_tree(w, P.T);
    w.put(P.n_tofar);
    w.put(P.n_fine);
    w.put(P.n_fine_uncovered);
    w.put(P.up_budget);
#define IFGF_W(name) w.vec(P.name);
    IFGF_PLAN_ARRAYS(IFGF_W)
#undef IFGF_W
    w.put(int64_t(P.up.size()));
    for (const UpLevel& u : P.up)
        walk_up(w, u);
    w.put(int64_t(P.up_parts.size()));
    for (const Plan::UpPart& u : P.up_parts)
    {
        w.put(u.level);
        walk_up(w, u.U);
    }
}

void walk_plan(Reader& r, Plan& P)
{
    double kr = 0, ki = 0;
    r.get(P.K.kind);
    r.get(kr);
    r.get(ki);
    P.K.k = cplx(kr, ki);
    r.get(P.K.kappa);
    r.get(P.K.dx);
    r.get(P.K.tol);
    r.get(P.base_els);
    r.get(P.base_order);
    r.get(P.ho_order);
    r.get(P.lo_order);
    r.get(P.factor);
    walk_tree(r, P.S);
    walk_tree(r, P.T);
    r.get(P.n_tofar);
    r.get(P.n_fine);
    r.get(P.n_fine_uncovered);
    r.get(P.up_budget);
#define IFGF_R(name) r.vec(P.name);
    IFGF_PLAN_ARRAYS(IFGF_R)
#undef IFGF_R
    // each level table and part holds at least 17 bytes: a count beyond the bytes left is
    // a damaged file, not an allocation to attempt
    int64_t n = -1;
    r.get(n);
    if (!r.ok || n < 0 || uint64_t(n) > r.left / 17)
    {
        r.ok = false;
        return;
    }
    P.up.resize(size_t(n));
    for (UpLevel& u : P.up)
        walk_up(r, u);
    n = -1;
    r.get(n);
    if (!r.ok || n < 0 || uint64_t(n) > r.left / 21)
    {
        r.ok = false;
        return;
    }
    P.up_parts.resize(size_t(n));
    for (Plan::UpPart& u : P.up_parts)
    {
        r.get(u.level);
        walk_up(r, u.U);
    }
}
#undef IFGF_PLAN_ARRAYS
} // namespace

std::string plan_save(const Plan& P, const PlanFileInfo& info, const char* path)
{
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return std::string("cannot open ") + path + " for writing";
    Writer w;
    w.f = f;
    bool ok = std::fwrite(MAGIC, 1, 8, f) == 8 && std::fwrite(&VERSION, 4, 1, f) == 1 &&
              std::fwrite(&ORDER_MARK, 4, 1, f) == 1;
    w.put(info);
    walk_plan(w, P);
    const uint64_t bytes = w.sum.bytes, hash = w.sum.done();
    ok = ok && w.ok && std::fwrite(&bytes, 8, 1, f) == 1 && std::fwrite(&hash, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
    {
        std::remove(path);
        return std::string("writing ") + path + " failed";
    }
    return "";
}

std::string plan_load(Plan& P, PlanFileInfo& info, const char* path)
{
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return std::string("cannot open ") + path;
    std::vector<unsigned char> buf;
    bool read_ok = std::fseek(f, 0, SEEK_END) == 0;
    const long size = read_ok ? std::ftell(f) : -1;
    read_ok = read_ok && size >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (read_ok)
    {
        buf.resize(size_t(size));
        read_ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    }
    std::fclose(f);
    if (!read_ok)
        return std::string("cannot read ") + path;
    const std::string bad = std::string(path) + " is not a plan file of this library";
    if (buf.size() < 16 + 16 || std::memcmp(buf.data(), MAGIC, 8) != 0)
        return bad;
    uint32_t version = 0, order = 0;
    std::memcpy(&version, buf.data() + 8, 4);
    std::memcpy(&order, buf.data() + 12, 4);
    if (order != ORDER_MARK)
        return bad + " (written on a machine of the other byte order)";
    if (version != VERSION)
        return bad + " (format version " + std::to_string(version) + ", this library reads " +
               std::to_string(VERSION) + ")";
    uint64_t bytes = 0, hash = 0;
    std::memcpy(&bytes, buf.data() + buf.size() - 16, 8);
    std::memcpy(&hash, buf.data() + buf.size() - 8, 8);
    const unsigned char* pay = buf.data() + 16;
    const size_t n_pay = buf.size() - 32;
    if (bytes != n_pay)
        return std::string(path) + " is truncated or damaged (payload length)";
    Checksum sum;
    sum.add(pay, n_pay);
    if (sum.done() != hash)
        return std::string(path) + " is damaged (checksum)";
    Reader r;
    r.p = pay;
    r.left = n_pay;
    r.get(info);
    Plan Q;
    walk_plan(r, Q);
    if (!r.ok || r.left != 0)
        return std::string(path) + " is damaged (array lengths)";
    P = std::move(Q);
    return "";
}

} // namespace ifgf

namespace ifgf
{
namespace
{
std::string tree_check(const Tree& t, const char* which)
{
    const std::string w = std::string(which) + " tree: ";
    const int64_t nb = t.num_boxes();
    if (t.n < 1 || int64_t(t.pts.size()) != 3 * t.n || int64_t(t.perm.size()) != t.n)
        return w + "point arrays";
    std::vector<uint8_t> seen(size_t(t.n), 0);
    for (int64_t p : t.perm)
    {
        if (p < 0 || p >= t.n || seen[size_t(p)])
            return w + "permutation";
        seen[size_t(p)] = 1;
    }
    if (t.levels < 1 || int64_t(t.level_off.size()) != t.levels + 1 || t.level_off[0] != 0 ||
        t.level_off[size_t(t.levels)] != nb)
        return w + "level offsets";
    for (int l = 0; l < t.levels; l++)
        if (t.level_off[size_t(l)] > t.level_off[size_t(l) + 1])
            return w + "level offsets";
    if (int64_t(t.end.size()) != nb || int64_t(t.parent.size()) != nb ||
        int64_t(t.leaf.size()) != nb || int64_t(t.side.size()) != nb ||
        int64_t(t.child.size()) != 8 * nb || int64_t(t.blo.size()) != 3 * nb ||
        int64_t(t.bhi.size()) != 3 * nb || int64_t(t.ctr.size()) != 3 * nb)
        return w + "box arrays";
    for (int64_t g = 0; g < nb; g++)
        if (t.begin[size_t(g)] < 0 || t.begin[size_t(g)] > t.end[size_t(g)] ||
            t.end[size_t(g)] > t.n || t.parent[size_t(g)] < -1 || t.parent[size_t(g)] >= nb)
            return w + "box ranges";
    for (int32_t c : t.child)
        if (c < -1 || c >= nb)
            return w + "children";
    return "";
}

template <typename T> bool monotone_to(const std::vector<T>& v, size_t n, int64_t last)
{
    if (v.size() != n || v.empty() || v[0] != 0 || int64_t(v.back()) != last)
        return false;
    for (size_t i = 1; i < v.size(); i++)
        if (v[i] < v[i - 1])
            return false;
    return true;
}
} // namespace

std::string plan_check(const Plan& P)
{
    std::string e = tree_check(P.S, "source");
    if (e.empty())
        e = tree_check(P.T, "target");
    if (!e.empty())
        return e;
    const int64_t nbs = P.S.num_boxes(), nbt = P.T.num_boxes(), nc = P.num_cones();
    if (!P.normals.empty() && int64_t(P.normals.size()) != 3 * P.S.n)
        return "normals";
    for (const std::vector<Pair>* L : {&P.far, &P.near})
        for (const Pair& p : *L)
            if (p.src < 0 || p.src >= nbs || p.tgt < 0 || p.tgt >= nbt)
                return "interaction lists";
    if (int64_t(P.has_far_anc.size()) != nbs || int64_t(P.parent_has_far.size()) != nbs ||
        int64_t(P.dom_lo.size()) != 3 * nbs || int64_t(P.dom_hi.size()) != 3 * nbs ||
        int64_t(P.nc.size()) != 3 * nbs || int64_t(P.nf.size()) != 3 * nbs ||
        int64_t(P.hc.size()) != 3 * nbs || int64_t(P.hf.size()) != 3 * nbs)
        return "cone grids";
    if (!monotone_to(P.table_off, size_t(nbs) + 1, int64_t(P.table.size())) ||
        !monotone_to(P.cone_off, size_t(nbs) + 1, nc))
        return "cone offsets";
    for (int64_t b = 0; b < nbs; b++)
    {
        int64_t cells = 1;
        for (int d = 0; d < 3; d++)
        {
            const int64_t n = P.nc[size_t(3 * b + d)];
            if (n < 0 || n > (int64_t(1) << 20))
                return "cone grids";
            cells *= n;
        }
        const int64_t slots = P.table_off[size_t(b) + 1] - P.table_off[size_t(b)];
        if (slots != 0 && slots != cells)
            return "cone tables";
    }
    if (int64_t(P.cone_cell.size()) != nc || int64_t(P.cone_mask.size()) != nc ||
        int64_t(P.cone_fine.size()) != nc)
        return "cone arrays";
    for (int64_t c = 0; c < nc; c++)
    {
        const int32_t b = P.cone_box[size_t(c)];
        if (b < 0 || b >= nbs || c < P.cone_off[size_t(b)] || c >= P.cone_off[size_t(b) + 1])
            return "cone boxes";
        if (P.cone_fine[size_t(c)] < 0 ||
            P.cone_fine[size_t(c)] + __builtin_popcount(P.cone_mask[size_t(c)]) > P.n_fine)
            return "fine slots";
    }
    for (int32_t v : P.table)
        if (v < -1 || v >= nc)
            return "cone tables";
    const int64_t n_hi = int64_t(P.ho_order[0]) * P.ho_order[1] * P.ho_order[2];
    auto up_check = [&](const UpLevel& u) {
        if (!u.valid)
            return true;
        int32_t compact = 0;
        for (int32_t v : u.cellmap)
        {
            if (v < -1)
                return false;
            compact = std::max(compact, v + 1);
        }
        if (int64_t(u.entries.size()) != int64_t(compact) * n_hi * 8)
            return false;
        for (const UpEntry& x : u.entries)
            if (x.fine < -1 || x.cf < -1 || x.sub < -1 || x.sub >= 27)
                return false;
        return true;
    };
    for (const UpLevel& u : P.up)
        if (!up_check(u))
            return "upward tables";
    for (const Plan::UpPart& u : P.up_parts)
        if (u.level < 0 || u.level >= P.S.levels || !up_check(u.U))
            return "upward tables";
    if (!P.box_up_part.empty())
    {
        if (int64_t(P.box_up_part.size()) != nbs || int64_t(P.box_k.size()) != 3 * nbs)
            return "geometry parts";
        for (int32_t v : P.box_up_part)
            if (v < -1 || v >= int32_t(P.up_parts.size()))
                return "geometry parts";
    }
    return "";
}
} // namespace ifgf
