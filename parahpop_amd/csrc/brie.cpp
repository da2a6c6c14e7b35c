// .brie reader (CBOR, layouts 1 and 2) and chain construction; see brie.h.
#include "brie.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace phx {
namespace {

struct Cbor {
    const uint8_t* p;
    const uint8_t* end;
    std::string file;

    [[noreturn]] void fail(const std::string& what) const { throw std::runtime_error(file + ": CBOR " + what); }
    uint8_t peek() const {
        if (p >= end) fail("truncated");
        return *p;
    }
    uint64_t arg(uint8_t info) {
        if (info < 24) return info;
        int bytes = info == 24 ? 1 : info == 25 ? 2 : info == 26 ? 4 : info == 27 ? 8 : 0;
        if (!bytes) fail("indefinite or reserved length not supported");
        if (end - p < bytes) fail("truncated");
        uint64_t v = 0;
        for (int i = 0; i < bytes; ++i) v = (v << 8) | *p++;
        return v;
    }
    // header of a definite-length item of major type `major`
    uint64_t head(int major) {
        const uint8_t b = peek();
        if ((b >> 5) != major) fail("unexpected item type");
        ++p;
        return arg(b & 31);
    }
    std::string text() {
        const uint64_t n = head(3);
        if (uint64_t(end - p) < n) fail("truncated text");
        std::string s(reinterpret_cast<const char*>(p), size_t(n));
        p += n;
        return s;
    }
    double number() {
        const uint8_t b = peek();
        const int major = b >> 5, info = b & 31;
        if (major == 0) return double(head(0));
        if (major == 1) return -1.0 - double(head(1));
        if (major != 7) fail("expected a number");
        ++p;
        if (info == 27) {
            if (end - p < 8) fail("truncated");
            uint64_t v = 0;
            for (int i = 0; i < 8; ++i) v = (v << 8) | *p++;
            double d;
            std::memcpy(&d, &v, 8);
            return d;
        }
        if (info == 26) {
            if (end - p < 4) fail("truncated");
            uint32_t v = 0;
            for (int i = 0; i < 4; ++i) v = (v << 8) | *p++;
            float f;
            std::memcpy(&f, &v, 4);
            return double(f);
        }
        if (info == 25) {
            if (end - p < 2) fail("truncated");
            const unsigned h = (unsigned(p[0]) << 8) | p[1];
            p += 2;
            const int e = (h >> 10) & 31, m = h & 1023;
            const double mag = e == 0 ? std::ldexp(double(m), -24) : e == 31 ? (m ? NAN : INFINITY)
                                                                              : std::ldexp(double(m + 1024), e - 25);
            return (h & 0x8000) ? -mag : mag;
        }
        fail("expected a number");
    }
    long long integer() {
        const double v = number();
        if (v != std::floor(v)) fail("expected an integer");
        return (long long)v;
    }
    void numbers(std::vector<double>& out) {
        const uint64_t n = head(4);
        out.reserve(out.size() + size_t(n));
        for (uint64_t i = 0; i < n; ++i) out.push_back(number());
    }
    void skip() {
        const uint8_t b = peek();
        const int major = b >> 5;
        if (major == 7) {
            ++p;
            const int info = b & 31;
            const int bytes = info == 24 ? 1 : info == 25 ? 2 : info == 26 ? 4 : info == 27 ? 8 : 0;
            if (end - p < bytes) fail("truncated");
            p += bytes;
            return;
        }
        ++p;
        const uint64_t n = arg(b & 31);
        if (major == 0 || major == 1) return;
        if (major == 2 || major == 3) {
            if (uint64_t(end - p) < n) fail("truncated");
            p += n;
        } else if (major == 4) {
            for (uint64_t i = 0; i < n; ++i) skip();
        } else if (major == 5) {
            for (uint64_t i = 0; i < 2 * n; ++i) skip();
        } else if (major == 6) {
            skip();
        }
    }
};

enum { F_FRAME, F_DTYPE, F_TARGET, F_CENTER, F_NINTERVALS, F_PDEG, F_DATAOFFSET, F_DATASIZE, F_INT };

void readLayout2(Cbor& c, BrieTable& t) {
    const uint64_t pairs = c.head(5);
    long long nUnits = -1;
    std::vector<double> ints, reals, data;
    for (uint64_t i = 0; i < pairs; ++i) {
        const std::string k = c.text();
        if (k == "metadata") {
            const uint64_t mp = c.head(5);
            for (uint64_t j = 0; j < mp; ++j) {
                const std::string mk = c.text();
                if (mk == "nBodyUnits") nUnits = c.integer();
                else if (mk == "intMetadata") c.numbers(ints);
                else if (mk == "doubleMetadata") c.numbers(reals);
                else c.skip();
            }
        } else if (k == "data") {
            c.numbers(data);
        } else {
            c.skip();
        }
    }
    if (nUnits <= 0 || ints.size() != size_t(F_INT * nUnits) || reals.size() != size_t(2 * nUnits))
        c.fail("layout-2 metadata sizes do not match nBodyUnits");
    const long long base = (long long)t.data.size();
    for (long long u = 0; u < nUnits; ++u) {
        BrieUnit b;
        auto f = [&](int field) { return (long long)ints[size_t(field * nUnits + u)]; };
        b.frame = int(f(F_FRAME));
        b.dtype = int(f(F_DTYPE));
        b.target = int(f(F_TARGET));
        b.center = int(f(F_CENTER));
        b.nIntervals = int(f(F_NINTERVALS));
        b.pdeg = int(f(F_PDEG));
        b.offset = base + f(F_DATAOFFSET);
        b.size = f(F_DATASIZE);
        b.start = reals[size_t(u)];
        b.final = reals[size_t(nUnits + u)];
        t.units.push_back(b);
    }
    t.data.insert(t.data.end(), data.begin(), data.end());
}

void readLayout1(Cbor& c, BrieTable& t) {
    const uint64_t n = c.head(4);
    for (uint64_t u = 0; u < n; ++u) {
        BrieUnit b;
        std::vector<double> data;
        const uint64_t pairs = c.head(5);
        for (uint64_t i = 0; i < pairs; ++i) {
            const std::string k = c.text();
            if (k == "metadata") {
                const uint64_t mp = c.head(5);
                for (uint64_t j = 0; j < mp; ++j) {
                    const std::string mk = c.text();
                    if (mk == "frame") b.frame = int(c.integer());
                    else if (mk == "dtype") b.dtype = int(c.integer());
                    else if (mk == "target") b.target = int(c.integer());
                    else if (mk == "center") b.center = int(c.integer());
                    else if (mk == "nintervals") b.nIntervals = int(c.integer());
                    else if (mk == "pdeg") b.pdeg = int(c.integer());
                    else if (mk == "startEpoch") b.start = c.number();
                    else if (mk == "finalEpoch") b.final = c.number();
                    else c.skip();
                }
            } else if (k == "data") {
                c.numbers(data);
            } else {
                c.skip();
            }
        }
        b.offset = (long long)t.data.size();
        b.size = (long long)data.size();
        t.data.insert(t.data.end(), data.begin(), data.end());
        t.units.push_back(b);
    }
}

}  // namespace

void loadBrie(const std::string& path, BrieTable& t) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to open file: " + path);
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Cbor c{bytes.data(), bytes.data() + bytes.size(), path};
    // single-element array wrappers around the top-level map
    while (c.p < c.end && (*c.p >> 5) == 4 && (*c.p & 31) == 1) ++c.p;
    const size_t first = t.units.size();
    const uint64_t pairs = c.head(5);
    int layout = 0;
    bool core = false;
    for (uint64_t i = 0; i < pairs; ++i) {
        const std::string k = c.text();
        if (k == "layout") {
            layout = int(c.integer());
        } else if (k == "core") {
            const int major = c.peek() >> 5;
            if (layout == 0) layout = major == 5 ? 2 : 1;
            if (layout == 2) readLayout2(c, t);
            else if (layout == 1) readLayout1(c, t);
            else c.fail("unsupported layout " + std::to_string(layout));
            core = true;
        } else {
            c.skip();
        }
    }
    if (!core) c.fail("no \"core\" entry");
    for (size_t u = first; u < t.units.size(); ++u) {
        const BrieUnit& b = t.units[u];
        const long long need = 1 + b.nIntervals + 3LL * b.nIntervals * (b.pdeg + 1);
        if (b.nIntervals <= 0 || b.pdeg < 0 || b.size < need || b.offset + b.size > (long long)t.data.size())
            throw std::runtime_error(path + ": body unit " + std::to_string(b.target) + " has inconsistent sizes");
        if (!(t.data[size_t(b.offset)] > 0.0))
            throw std::runtime_error(path + ": body unit " + std::to_string(b.target) + " has no positive interval radius");
    }
}

// Orientation metadata: 12 integer fields per unit (UNITTYPE, TARGET, eight
// unit-specific fields, DATAOFFSET, DATASIZE) and the envelope
// (orientations/metadata/Layout.h:12-98).
void loadBrot2(const std::string& path, BrotTable& t) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to open file: " + path);
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Cbor c{bytes.data(), bytes.data() + bytes.size(), path};
    while (c.p < c.end && (*c.p >> 5) == 4 && (*c.p & 31) == 1) ++c.p;
    const uint64_t pairs = c.head(5);
    int layout = 0;
    long long nUnits = -1;
    std::vector<double> ints, reals, data;
    for (uint64_t i = 0; i < pairs; ++i) {
        const std::string k = c.text();
        if (k == "layout") {
            layout = int(c.integer());
        } else if (k == "core") {
            if (layout != 2) c.fail("only layout-2 orientation files are supported");
            const uint64_t cp = c.head(5);
            for (uint64_t j = 0; j < cp; ++j) {
                const std::string ck = c.text();
                if (ck == "metadata") {
                    const uint64_t mp = c.head(5);
                    for (uint64_t q = 0; q < mp; ++q) {
                        const std::string mk = c.text();
                        if (mk == "nBodyUnits") nUnits = c.integer();
                        else if (mk == "intMetadata") c.numbers(ints);
                        else if (mk == "doubleMetadata") c.numbers(reals);
                        else c.skip();
                    }
                } else if (ck == "data") {
                    c.numbers(data);
                } else {
                    c.skip();
                }
            }
        } else {
            c.skip();
        }
    }
    constexpr int kInt = 12;
    if (nUnits <= 0 || ints.size() != size_t(kInt * nUnits) || reals.size() != size_t(2 * nUnits))
        c.fail("orientation metadata sizes do not match nBodyUnits");
    const long long base = (long long)t.data.size();
    size_t found = 0;
    for (long long u = 0; u < nUnits; ++u) {
        auto f = [&](int field) { return (long long)ints[size_t(field * nUnits + u)]; };
        if (f(0) >= 1 && f(0) <= 3) {
            BrotBody b;
            b.type = int(f(0));
            b.target = int(f(1));
            for (int k = 0; k < 8; ++k) b.p[k] = int(f(2 + k));
            b.offset = base + f(10);
            b.size = f(11);
            b.start = reals[size_t(u)];
            b.final = reals[size_t(nUnits + u)];
            long long need = 0;
            if (b.type == 1) {
                for (int k = 0; k < 3; ++k)
                    if (b.p[k] < 0) c.fail("body orientation unit with a negative polynomial degree");
                for (int k = 5; k < 8; ++k)
                    if (b.p[k] < 0) c.fail("body orientation unit with a negative term count");
                need = (long long)b.p[0] + b.p[1] + b.p[2] + 3 + b.p[5] + b.p[6] + b.p[7];
            } else if (b.type == 2) {
                if (b.p[0] < 0 || b.p[1] < 0) c.fail("nutation-precession unit with negative sizes");
                need = (long long)(b.p[0] + 1) * b.p[1];
            } else {
                if (b.p[2] < 1 || b.p[3] < 0) c.fail("binary frame unit without intervals");
                need = 1 + (long long)b.p[2] + 3LL * (b.p[3] + 1) * b.p[2];
            }
            if (f(10) < 0 || b.size < need || f(10) + b.size > (long long)data.size())
                c.fail("orientation unit with inconsistent sizes");
            t.bodies.push_back(b);
            ++found;
            continue;
        }
        if (f(0) != 4 || f(2) != 0) continue;   // unit_type 4, model 0 (IERS2000)
        BrotIers b;
        b.target = int(f(1));
        b.nNut = int(f(4));
        b.nErp = int(f(5));
        if (f(6) != 5 || f(7) != 3) c.fail("IERS2000 unit with Lagrange degrees other than 5 and 3");
        b.offset = base + f(10);
        const long long size = f(11);
        b.start = reals[size_t(u)];
        b.final = reals[size_t(nUnits + u)];
        const long long need = 3LL * b.nNut + 5LL * b.nErp + 4 + b.nErp;
        if (b.nNut < 6 || b.nErp < 4 || size < need || f(10) + size > (long long)data.size())
            c.fail("IERS2000 unit with inconsistent sizes");
        t.iers.push_back(b);
        ++found;
    }
    t.data.insert(t.data.end(), data.begin(), data.end());
    if (found == 0) throw std::runtime_error(path + ": no IERS2000 or body orientation unit");
}

namespace {

int unitOf(const BrieTable& t, int naif) {
    for (size_t u = 0; u < t.units.size(); ++u)
        if (t.units[u].target == naif) return int(u);
    return -1;
}

// naif ids from `naif` up to the root (inclusive), and the unit of each link
void ancestry(const BrieTable& t, int naif, std::vector<int>& ids, std::vector<int>& links) {
    ids.assign(1, naif);
    links.clear();
    for (int guard = 0; guard < 64; ++guard) {
        const int u = unitOf(t, ids.back());
        if (u < 0) return;
        links.push_back(u);
        ids.push_back(t.units[size_t(u)].center);
    }
    throw std::runtime_error("ephemeris: cyclic center chain at body " + std::to_string(naif));
}

}  // namespace

std::vector<BrieStep> brieChain(const BrieTable& t, int naif, int center) {
    std::vector<int> ia, la, ib, lb;
    ancestry(t, naif, ia, la);
    ancestry(t, center, ib, lb);
    // the highest common center: the first id of the target's ancestry that the
    // center's ancestry also holds
    int common = -1;
    size_t ka = 0, kb = 0;
    for (size_t i = 0; i < ia.size() && common < 0; ++i)
        for (size_t j = 0; j < ib.size(); ++j)
            if (ia[i] == ib[j]) {
                common = ia[i];
                ka = i;
                kb = j;
                break;
            }
    if (common < 0)
        throw std::runtime_error("ephemeris: no chain connects body " + std::to_string(naif) + " to body " +
                                 std::to_string(center));
    std::vector<BrieStep> out;
    for (size_t i = 0; i < ka; ++i) out.push_back({la[i], 1.0});
    for (size_t j = 0; j < kb; ++j) out.push_back({lb[j], -1.0});
    return out;
}

void brieCoverage(const BrieTable& t, const std::vector<BrieStep>& chain, double& start, double& end) {
    start = -INFINITY;
    end = INFINITY;
    for (const auto& s : chain) {
        const BrieUnit& u = t.units[size_t(s.unit)];
        const double* d = t.data.data() + u.offset;
        start = std::max(start, d[1]);
        end = std::min(end, d[1] + 2.0 * d[0] * u.nIntervals);
    }
}

}  // namespace phx
