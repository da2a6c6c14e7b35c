// C-ABI of libphx: context (configuration, environment tables on the device),
// device-resident batches and the propagation driver loop.
//
// Reference behaviour kept at this boundary (src/common/ unless noted):
//   target validation and unit conversion      driver.cpp:99-123
//   epoch conversion, first/last step sizes    engine/paraHPOP/propagators/FixedStep.h:7-47
//   end-of-simulation check                    driver.cpp:146-155
//   device budget guard (60 % of free memory)  engine/paraHPOP/model/environment/Environment.h:977-1050
// Compiled with -ffp-contract=off: the epoch arithmetic here must round
// exactly as the CPU reference does.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "phx_internal.h"
#include "phx_serve.h"

using namespace phx;

namespace {

std::string libraryDir() {
    Dl_info info;
    if (dladdr((void*)&phx_create, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t k = p.find_last_of('/');
        return k == std::string::npos ? "." : p.substr(0, k);
    }
    return ".";
}

std::string dirName(const std::string& p) {
    const size_t k = p.find_last_of('/');
    if (k == std::string::npos) return ".";
    return k == 0 ? "/" : p.substr(0, k);
}

}  // namespace

namespace phx {
const double kRkC[12] = {2. / 27, 1. / 9, 1. / 6, 5. / 12, 1. / 2, 5. / 6, 1. / 6, 2. / 3, 1. / 3, 1.0, 0, 1.0};
}  // namespace phx

namespace {

// TDB - TT (s) at a TT MJD: the series of EarthCorrections.h:46-56.
double tdbMinusTtHost(double mjdTt) {
    const double t = (mjdTt - 51544.5) / 36525.0;
    return 0.001657 * std::sin(628.3076 * t + 6.2401) + 0.000022 * std::sin(575.3385 * t + 4.2970) +
           0.000014 * std::sin(1256.6152 * t + 6.1969) + 0.000005 * std::sin(606.9777 * t + 4.0212) +
           0.000005 * std::sin(52.9691 * t + 0.4444) + 0.000002 * std::sin(21.3299 * t + 5.5431) +
           0.000010 * std::sin(628.3076 * t + 4.2490);
}

// TAI-UTC (s) that the EOP time model's ET -> UTC iteration settles on at `et`
// (EarthCorrections.h:87-105, the same steps as the device eopTimes): UTC =
// TT - 32.184 s - this value.
double utcConvention(const EopTable& e, double et) {
    double ttEt = et;
    for (int j = 0; j < 2; ++j) ttEt = et - tdbMinusTtHost(51544.5 + ttEt / 86400.0);
    double utc = 51544.5 + (ttEt - 69.184) / 86400.0;
    double dat = 0.0;
    for (int j = 0; j < 3; ++j) {
        double day = std::floor(utc - e.firstMjdUtc);
        day = std::fmin(std::fmax(day, 0.0), double(e.days - 1));
        dat = e.dat[size_t(day)];
        utc = 51544.5 + (ttEt - 32.184 - dat) / 86400.0;
    }
    return dat;
}

// The UT1 steps of the analytic orientation at the table's leap seconds. The
// time model's UTC steps back by each leap second while the table's UT1-UTC is
// interpolated without a step, so UT1 and the Earth rotation angle step with
// it. Each step of at least half a second (the leap seconds; the fractional
// TAI-UTC drift of tables before 1972 is left to the interpolation) is located
// to 1e-9 s where utcConvention changes, scanning the seconds around the
// table's TAI-UTC change; the offset after it is -(dERA/dUT1) times the sum of
// the steps so far. Empty without an EOP table or with a .brot2 orientation
// (whose UT1 - TT has no steps).
void computeLeaps(phx_ctx& c) {
    c.leapCount = 0;
    if (!c.env.hasEop || !c.env.brot.empty()) return;
    const EopTable& e = c.env.eop;
    const double rate = 2.0 * 3.141592653589793238462643383279502884 * 1.00273781191135448 / 86400.0;
    std::vector<double> at, theta;
    double stepped = 0.0;
    for (int i = 1; i < e.days; ++i) {
        if (!(std::fabs(e.dat[size_t(i)] - e.dat[size_t(i - 1)]) >= 0.5)) continue;
        const double mid = (e.firstMjdUtc + i - 51544.5) * 86400.0 + 32.184 + 0.5 * (e.dat[size_t(i)] + e.dat[size_t(i - 1)]);
        const double lo0 = mid - 0.5 * std::fabs(e.dat[size_t(i)] - e.dat[size_t(i - 1)]) - 3.0;
        const int n = 6000;   // 1 ms scan of the window, then bisection
        const double span = std::fabs(e.dat[size_t(i)] - e.dat[size_t(i - 1)]) + 6.0;
        double prev = utcConvention(e, lo0), prevT = lo0;
        for (int k = 1; k <= n; ++k) {
            const double t = lo0 + span * double(k) / n;
            const double v = utcConvention(e, t);
            if (v != prev) {
                double a = prevT, b = t;
                for (int it = 0; it < 60 && b - a > 1e-9; ++it) {
                    const double m = 0.5 * (a + b);
                    (utcConvention(e, m) == prev ? a : b) = m;
                }
                stepped += v - prev;
                at.push_back(b);
                theta.push_back(-rate * stepped);
                prev = v;
            }
            prevT = t;
        }
    }
    if (at.empty()) return;
    if (at.size() > 256) throw std::runtime_error("EOP table: more than 256 leap-second steps");
    std::vector<double> v(at);
    v.insert(v.end(), theta.begin(), theta.end());
    c.leaps.upload(v.data(), v.size(), c.stream);
    c.leapCount = int(at.size());
}

// The coefficients the NRLMSISE-00 chains read one at a time, gathered behind
// the table in the kernels' HOT_* layout: pdl, ptm and pdm whole (contiguous in
// the table), pt[0..14], ps[0], pd[k][0], ptl[k][0], pma[8][0].
void appendNrlHot(std::vector<double>& v) {
    // table offsets of the blocks (the reference's coefficient file order)
    constexpr int pt = 0, pd = 150, ps = 1500, pdl = 1650, ptl = 1790, pma = 2190;
    std::vector<double> h(v.begin() + pdl, v.begin() + pdl + 140);
    for (int i = 0; i < 15; ++i) h.push_back(v[size_t(pt + i)]);
    h.push_back(v[size_t(ps)]);
    for (int k = 0; k < 9; ++k) h.push_back(v[size_t(pd + 150 * k)]);
    for (int k = 0; k < 4; ++k) h.push_back(v[size_t(ptl + 100 * k)]);
    h.push_back(v[size_t(pma + 100 * 8)]);
    if (h.size() != 170) throw std::logic_error("NRLMSISE-00 gathered coefficient block");
    v.insert(v.end(), h.begin(), h.end());
}

void loadTables(phx_ctx& c) {
    const std::string dir = c.dataDir;
    if (c.env.model.tidesActive()) {
        std::ifstream in(dir + "/iers2010_solid_tide_frequency.txt");
        if (!in) throw std::runtime_error("missing table " + dir + "/iers2010_solid_tide_frequency.txt");
        std::vector<double> rows[3];
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream s(line);
            int order, mult[5];
            double ip, op;
            s >> order >> mult[0] >> mult[1] >> mult[2] >> mult[3] >> mult[4] >> ip >> op;
            if (!s || order < 0 || order > 2) throw std::runtime_error("bad tide frequency table row");
            for (int k = 0; k < 5; ++k) rows[order].push_back(mult[k]);
            rows[order].push_back(ip);
            rows[order].push_back(op);
        }
        std::vector<double> all;
        for (int o = 0; o < 3; ++o) {
            c.freqCount[o] = int(rows[o].size() / 7);
            all.insert(all.end(), rows[o].begin(), rows[o].end());
        }
        c.freq.upload(all.data(), all.size(), c.stream);
        std::ifstream f(dir + "/fes2004_ocean_tide.txt");
        if (!f) throw std::runtime_error("missing table " + dir + "/fes2004_ocean_tide.txt");
        std::vector<double> w, t;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream s(line);
            char tag;
            s >> tag;
            if (tag == 'W') {
                double v[8];
                for (double& x : v) s >> x;
                w.insert(w.end(), v, v + 8);
            } else if (tag == 'T') {
                double v[6];
                for (double& x : v) s >> x;
                t.insert(t.end(), v, v + 6);
            }
            if (!s) throw std::runtime_error("bad FES2004 table row");
        }
        c.waveCount = int(w.size() / 8);
        c.waves.upload(w.data(), w.size(), c.stream);
        c.terms.upload(t.data(), t.size(), c.stream);
    }
    if (c.env.hasWeather) {
        std::ifstream in(dir + "/nrlmsise00_coefficients.txt");
        if (!in) throw std::runtime_error("missing table " + dir + "/nrlmsise00_coefficients.txt");
        // blocks in file order: pt pd ps pdl ptm pdm ptl pma sam pavgm (3300 values)
        std::vector<double> v;
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#' || std::isalpha((unsigned char)line[0])) continue;
            std::istringstream s(line);
            double x;
            while (s >> x) v.push_back(x);
        }
        if (v.size() != 3300) throw std::runtime_error("NRLMSISE-00 coefficient table must hold 3300 values");
        if (uploadGlobeTrig(v.data()) != 0) throw HipError("globe phase upload failed");
        appendNrlHot(v);
        c.nrl.upload(v.data(), v.size(), c.stream);
        const WeatherTable& wt = c.env.weather;
        c.weather.upload(wt.data.data(), wt.data.size(), c.stream);
    }
    if (c.env.hasEop) {
        const EopTable& e = c.env.eop;
        std::vector<double> v;
        v.insert(v.end(), e.xp.begin(), e.xp.end());
        v.insert(v.end(), e.yp.begin(), e.yp.end());
        v.insert(v.end(), e.dut.begin(), e.dut.end());
        v.insert(v.end(), e.dat.begin(), e.dat.end());
        c.eop.upload(v.data(), v.size(), c.stream);
    }
}

// Packed SH term stream (see ShTerm in phx_common.h), its order starts and the
// diagonal factors followed by their cumulative products.
void packShField(const ShField& f, std::vector<double>& terms, std::vector<float>& terms32, std::vector<int>& start,
                 std::vector<double>& diag) {
    const int L = f.degree, M = f.order;
    terms.clear();
    terms32.clear();
    start.assign(static_cast<size_t>(L) + 2, 0);
    diag.assign(static_cast<size_t>(L) + 1, 0.0);
    int k = 0;
    for (int m = 0; m <= L; ++m) {
        diag[size_t(m)] = f.factors[2 * f.index(m, m)];
        start[size_t(m)] = k;
        if (m > M) continue;
        // each order's run is padded with zero terms to a multiple of four
        // (the kernel's FP64 loop reads four-term batches)
        for (int n = m; n < m + ((L - m + 4) & ~3); ++n, ++k) {
            if (n > L) {
                for (int p = 0; p < SHT_STRIDE; ++p) {
                    terms.push_back(0.0);
                    terms32.push_back(0.0f);
                }
                continue;
            }
            const double C = f.C[f.index(n, m)], S = m > 0 ? f.S[f.index(n, m)] : 0.0;
            double e;
            if (f.normalized) e = n == m ? 0.0 : std::sqrt((2.0 * n + 1.0) * (double(n) * n - double(m) * m) / (2.0 * n - 1.0));
            else e = double(n + m);
            double a = 0.0, b = 0.0;
            if (n < L) {
                a = f.factors[2 * f.index(n + 1, m)];
                b = f.factors[2 * f.index(n + 1, m) + 1];
            }
            const double t[SHT_STRIDE] = {(n + 1) * C, (n + 1) * S, C, S, e * C, e * S, a, b};
            terms.insert(terms.end(), t, t + SHT_STRIDE);
            for (double x : t) terms32.push_back(float(x));
        }
    }
    start[size_t(L) + 1] = k;
    // one zero term of padding: the pipelined loop reads one term ahead
    for (int p = 0; p < SHT_STRIDE; ++p) {
        terms.push_back(0.0);
        terms32.push_back(0.0f);
    }
    // then the cumulative products prod_{k=1..m} c_k (the seed of an order block)
    double prod = 1.0;
    diag.push_back(1.0);
    for (int m = 1; m <= L; ++m) {
        prod *= diag[size_t(m)];
        diag.push_back(prod);
    }
}

void buildShTables(phx_ctx& c) {
    std::vector<double> terms, diag;
    std::vector<float> terms32;
    std::vector<int> start;
    packShField(c.env.sh, terms, terms32, start, diag);
    c.shTerms.upload(terms.data(), terms.size(), c.stream);
    c.shTerms32.upload(terms32.data(), terms32.size(), c.stream);
    c.shOrderStart.upload(start.data(), start.size(), c.stream);
    c.shDiag.upload(diag.data(), diag.size(), c.stream);
}

// Term streams of the fields of the bodies other than the Earth and the
// orientation table of their axes; fills the device pointers of ForceModel.
void buildRotTables(phx_ctx& c) {
    if (c.env.rot.empty()) return;
    c.brotBody.upload(c.env.brot.data.data(), c.env.brot.data.size(), c.stream);
    for (size_t k = 0; k < c.env.rot.size(); ++k) {
        const RotBody& r = c.env.rot[k];
        if (!r.sh) continue;
        std::vector<double> terms, diag;
        std::vector<float> terms32;
        std::vector<int> start;
        packShField(r.field, terms, terms32, start, diag);
        c.rotTerms[k].upload(terms.data(), terms.size(), c.stream);
        c.rotStart[k].upload(start.data(), start.size(), c.stream);
        c.rotDiag[k].upload(diag.data(), diag.size(), c.stream);
        c.fm.rotTerms[k] = c.rotTerms[k].p;
        c.fm.rotOrderStart[k] = c.rotStart[k].p;
        c.fm.rotDiag[k] = c.rotDiag[k].p;
    }
}

// Body of ephemeris slot `slot`: the configured bodies in order, then the Sun
// and the Moon of the tide model (-1: unused slot).
int ephSlotBody(const ModelConfig& m, int slot) {
    if (slot < int(m.bodies.size())) return m.bodies[size_t(slot)].naif;
    if (slot == kMaxBodies && m.tidesActive()) return NAIF_SUN;
    if (slot == kMaxBodies + 1 && m.tidesActive()) return NAIF_MOON;
    return -1;
}

void buildForceModel(phx_ctx& c) {
    const ModelConfig& m = c.env.model;
    ForceModel& fm = c.fm;
    std::memset(&fm, 0, sizeof fm);
    if (m.bodies.size() > size_t(kMaxBodies)) throw std::runtime_error("too many bodies");
    fm.nBodies = int(m.bodies.size());
    int occ = 0;
    for (size_t i = 0; i < m.bodies.size(); ++i) {
        const BodySpec& b = m.bodies[i];
        fm.bodyNaif[i] = b.naif;
        fm.bodyGm[i] = b.gm;
        fm.bodyPointGravity[i] = b.pointGravity ? 1 : 0;
        if (b.naif == NAIF_EARTH) {
            fm.earthPointGravity = b.pointGravity ? 1 : 0;
            fm.gmEarth = b.gm;
            if (b.atmosphere && b.nrlmsise) {
                fm.drag = 1;
                fm.hCutoff = b.nrl.hCutoff;
                fm.densityScale = b.nrl.densityScale;
                fm.earthRadius = b.radius;
                fm.flattening = b.flattening;
            } else if (b.atmosphere && !b.expSegments.empty()) {
                if (b.expSegments.size() > size_t(kMaxExpSegments))
                    throw std::runtime_error("at most " + std::to_string(kMaxExpSegments) + " exponential atmosphere bands are supported");
                fm.expDrag = 1;
                fm.nExpSeg = int(b.expSegments.size());
                for (size_t s = 0; s < b.expSegments.size(); ++s) {
                    fm.expSeg[s][0] = b.expSegments[s].rho0;
                    fm.expSeg[s][1] = b.expSegments[s].h0;
                    fm.expSeg[s][2] = b.expSegments[s].scaleHeight;
                    fm.expSeg[s][3] = b.expSegments[s].hCutoff;
                }
                fm.earthRadius = b.radius;
                fm.flattening = b.flattening;
            }
            if (b.oblateness && !b.shape) {
                fm.hasJ2 = 1;
                fm.j2Gm = b.gm;
                fm.j2Coef = b.j2;
                fm.j2Radius = b.radius;
                fm.j2Soi = b.soi;
            }
        }
        if (b.radiation) {
            fm.srp = 1;
            fm.srpBody = int(i);
            fm.srpFactor = m.au * m.au * 4.56e-3;
            fm.sunRadius = b.radius;
        }
        if (b.occulting && b.naif != NAIF_SUN) {
            if (occ >= 4) throw std::runtime_error("at most four occulting bodies are supported");
            fm.occBody[occ] = int(i);
            fm.occRadius[occ] = b.radius;
            ++occ;
        }
    }
    fm.nOccluders = occ;
    // fields of the bodies other than the Earth (device pointers: buildRotTables)
    fm.nRot = int(c.env.rot.size());
    for (size_t k = 0; k < c.env.rot.size(); ++k) {
        const RotBody& r = c.env.rot[k];
        const BodySpec& b = m.bodies[size_t(r.body)];
        fm.rotBody[k] = r.body;
        fm.rotSh[k] = r.sh ? 1 : 0;
        fm.rotSoi[k] = b.soi;
        if (r.sh) {
            fm.rotDegree[k] = r.field.degree;
            fm.rotOrder[k] = r.field.order;
            fm.rotNormalized[k] = r.field.normalized ? 1 : 0;
            fm.rotGm[k] = r.field.gm;
            fm.rotRadius[k] = r.field.radius;
        } else {
            fm.rotGm[k] = b.gm;
            fm.rotRadius[k] = b.radius;
            fm.rotJ2[k] = b.j2;
        }
    }
    if (c.env.hasSh) {
        fm.hasSh = 1;
        fm.shDegree = c.env.sh.degree;
        fm.shOrder = c.env.sh.order;
        fm.shRadius = c.env.sh.radius;
        fm.shGm = c.env.sh.gm;
        fm.shNormalized = c.env.sh.normalized ? 1 : 0;
    }
    fm.tides = m.tidesActive() ? 1 : 0;
    // Chebyshev ephemeris: chains of the body slots (bodies, then Sun and Moon)
    c.ephChains.clear();
    if (!c.env.brie.empty()) {
        for (int slot = 0; slot < kEphSlots; ++slot) {
            const int naif = ephSlotBody(m, slot);
            if (naif < 0) continue;
            auto ch = brieChain(c.env.brie, naif, NAIF_EARTH);
            if (ch.size() > size_t(kEphSteps))
                throw std::runtime_error("ephemeris chain of body " + std::to_string(naif) + " is longer than " +
                                         std::to_string(kEphSteps) + " links");
            c.ephChains.push_back(ch);
        }
    }
    fm.relativity = m.relativity ? 1 : 0;
    fm.muEarthSI = 398600.4415e9;
    fm.radiusSI = 6378.1363e3;
    fm.clightSI = 299792457.999999984;
}

}  // namespace

namespace phx {

EpochArgs epochArgs(const phx_ctx& c) {
    EpochArgs a{};
    a.eop = c.eop.p;
    a.eopDays = c.env.hasEop ? c.env.eop.days : 0;
    a.eopFirst = c.env.hasEop ? c.env.eop.firstMjdUtc : 0.0;
    a.solid = c.env.model.solidTides ? 1 : 0;
    a.ocean = c.env.model.oceanTides ? 1 : 0;
    a.zeroTide = c.env.model.zeroTide ? 1 : 0;
    a.freq = c.freq.p;
    for (int k = 0; k < 3; ++k) a.freqCount[k] = c.freqCount[k];
    a.waves = c.waves.p;
    a.waveCount = c.waveCount;
    a.terms = c.terms.p;
    a.nrlCoeff = c.nrl.p;
    a.weather = c.weather.p;
    a.wDays = c.env.hasWeather ? c.env.weather.days : 0;
    a.wStart = c.env.hasWeather ? c.env.weather.startMjdUtc : 0.0;
    a.eph = c.eph.p;
    a.ephChain = c.ephChain.p;
    a.iers = c.iers.p;
    a.iersNut = c.env.brot.empty() ? 0 : c.env.brot.iers[0].nNut;
    a.iersErp = c.env.brot.empty() ? 0 : c.env.brot.iers[0].nErp;
    a.leaps = c.leaps.p;
    a.leapCount = c.leapCount;
    a.gridRows = 0;
    a.brot = c.brotBody.p;
    a.nRot = int(c.env.rot.size());
    for (size_t k = 0; k < c.env.rot.size(); ++k) {
        const RotBody& r = c.env.rot[k];
        const BrotBody& u = c.env.brot.bodies[size_t(r.unit)];
        RotUnit& d = a.rot[k];
        d.type = u.type;
        d.off = u.offset;
        d.bary = -1;
        if (u.type == 1) {
            d.degRa = u.p[0];
            d.degDec = u.p[1];
            d.degPm = u.p[2];
            d.npRa = u.p[5];
            d.npDec = u.p[6];
            d.npPm = u.p[7];
            if (r.bary >= 0) {
                const BrotBody& bu = c.env.brot.bodies[size_t(r.bary)];
                d.bary = bu.offset;
                d.baryDeg = bu.p[0];
                d.baryAngles = bu.p[1];
            }
        } else {
            d.inertFrame = u.p[0];
            d.nIntervals = u.p[2];
            d.pdeg = u.p[3];
        }
    }
    return a;
}

bool needsCoverageCheck(const phx_ctx& c) { return !c.ephChains.empty() || !c.env.brot.empty() || !c.env.rot.empty(); }

void checkEphemerisCoverage(const phx_ctx& c, double lo, double hi) {
    // binary frames of the body fields carry an envelope (model-level coverage,
    // interface/config/model/environment/Orientations.h:150-200)
    for (const RotBody& r : c.env.rot) {
        const BrotBody& u = c.env.brot.bodies[size_t(r.unit)];
        if (u.type == 3 && (lo < u.start || hi > u.final)) {
            std::ostringstream m;
            m.precision(17);
            m << "Epoch range [" << lo << ", " << hi << "] not in the available orientation range [" << u.start << ", "
              << u.final << "] of rotation target " << u.target;
            throw std::runtime_error(m.str());
        }
    }
    if (!c.env.brot.empty()) {
        const BrotIers& u = c.env.brot.iers[0];
        if (lo < u.start || hi > u.final) {
            std::ostringstream m;
            m.precision(17);
            m << "Epoch range [" << lo << ", " << hi << "] not in the available orientation range [" << u.start << ", "
              << u.final << "]";
            throw std::runtime_error(m.str());
        }
    }
    for (const auto& ch : c.ephChains) {
        double s0, s1;
        brieCoverage(c.env.brie, ch, s0, s1);
        if (lo < s0 || hi > s1) {
            std::ostringstream m;
            m.precision(17);
            m << "Epoch range [" << lo << ", " << hi << "] not in the available ephemeris range [" << s0 << ", " << s1 << "]";
            throw std::runtime_error(m.str());
        }
    }
}

void splitOrderBlocks(const phx_ctx& c, int W, int G, int* out) {
    // split SH kernel: costs in units of one FP64 term of one wave. Order m
    // runs its padded term count (order 0 with its derivative recursion
    // costs more per term) plus a fixed per-order cost; the waves that carry
    // point gravity (W - 1), tides and relativity (W - 2), SRP (W - 3) and
    // the NRLMSISE-00 lane inputs (W - 4) of the first workgroup start with
    // that cost. Every wave gets the same total where it can (water filling).
    const int L = c.fm.shDegree, M = c.fm.shOrder;
    const int Wt = W * G;
    std::vector<double> cost(static_cast<size_t>(M) + 1);
    double sum = 0.0;
    for (int m = 0; m <= M; ++m) {
        cost[size_t(m)] = double((L - m + 4) & ~3) * (m == 0 ? c.splitCostOrder0 : 1.0) + c.splitCostOrder;
        sum += cost[size_t(m)];
    }
    std::vector<double> extra(static_cast<size_t>(Wt), 0.0);
    extra[size_t(W - 1)] = c.splitCostPg;
    if (c.fm.tides || c.fm.relativity) extra[size_t(W - 2)] = c.splitCostCorr;
    if (c.fm.srp) extra[size_t(W - 3)] = c.splitCostSrp;
    if (c.fm.drag) extra[size_t(W - 4)] = c.splitCostLane;
    // level T with sum_w max(0, T - extra_w) = sum (bisection)
    double lo = 0.0, hi = sum + *std::max_element(extra.begin(), extra.end());
    for (int it = 0; it < 100; ++it) {
        const double T = 0.5 * (lo + hi);
        double got = 0.0;
        for (double e : extra) got += std::max(0.0, T - e);
        (got < sum ? lo : hi) = T;
    }
    const double T = hi;
    out[0] = 0;
    int m = 0;
    double acc = 0.0, target = 0.0;
    for (int w = 0; w < Wt; ++w) {
        target += std::max(0.0, T - extra[size_t(w)]);
        if (w == Wt - 1) {
            m = M + 1;
        } else {
            while (m <= M && acc + 0.5 * cost[size_t(m)] <= target) {
                acc += cost[size_t(m)];
                ++m;
            }
        }
        out[w + 1] = m;
    }
}

void orderBlocks(const phx_ctx& c, int W, int out[kMaxWaves + 1], bool split) {
    const int L = c.fm.shDegree, M = c.fm.shOrder;
    if (split && W >= 8) {
        splitOrderBlocks(c, W, 1, out);
        for (int w = W + 1; w <= kMaxWaves; ++w) out[w] = out[W];
        return;
    }
    std::vector<double> share(static_cast<size_t>(W), 1.0);
    const int corrW = W >= 3 ? W - 2 : W - 1;
    if (W > 1) {
        share[0] *= c.share0;
        if (c.fm.tides || c.fm.relativity || c.fm.srp) share[size_t(corrW)] *= c.shareCorr;
    }
    double total = 0.0;
    for (int m = 0; m <= M; ++m) total += double(L - m + 1) + 0.6 * m / std::max(1, W);
    double sumShare = 0.0;
    for (double s : share) sumShare += s;
    out[0] = 0;
    int m = 0;
    double acc = 0.0, target = 0.0;
    for (int w = 0; w < W; ++w) {
        target += total * share[size_t(w)] / sumShare;
        if (w == W - 1) {
            m = M + 1;
        } else {
            while (m <= M && acc + 0.5 * (double(L - m + 1)) <= target) {
                acc += double(L - m + 1) + 0.6 * m / std::max(1, W);
                ++m;
            }
        }
        out[w + 1] = m;
    }
    for (int w = W + 1; w <= kMaxWaves; ++w) out[w] = out[W];
}

void validateTargets(size_t n, const int64_t* id, const double* epoch, const double* state6, const double* mass,
                     const double* drag, const double* srp, const double* cd, const double* cr) {
    if (n == 0) throw std::invalid_argument("targets must be a nonempty array.");
    if (n > size_t(0x7fffffff)) throw std::invalid_argument("Too many targets for the engine index type.");
    std::set<int64_t> ids;
    for (size_t i = 0; i < n; ++i) {
        if (!ids.insert(id[i]).second) throw std::invalid_argument("Target IDs must be unique.");
        for (int d = 0; d < 6; ++d)
            if (!std::isfinite(state6[d * n + i])) throw std::invalid_argument("State contains a nonfinite value.");
        if (!std::isfinite(epoch[i]) || !std::isfinite(mass[i]) || !std::isfinite(drag[i]) || !std::isfinite(srp[i]) ||
            !std::isfinite(cd[i]) || !std::isfinite(cr[i]) || mass[i] <= 0 || drag[i] <= 0 || srp[i] < 0 || cd[i] < 0 ||
            cr[i] < 0)
            throw std::invalid_argument("Invalid target epoch or physical parameter.");
    }
}

}  // namespace phx

namespace {

int fail(phx_ctx* c, const std::exception& e) {
    if (c) c->setError(e.what());
    if (dynamic_cast<const HipError*>(&e)) return PHX_EDEVICE;
    if (dynamic_cast<const std::bad_alloc*>(&e)) return PHX_ENOMEM;
    return PHX_EINVAL;
}

thread_local std::string g_createError;

}  // namespace

namespace phx {

// Builds the device batch: validation, unit conversion, epoch groups, tiles,
// per-group step sequence and epoch records.
std::unique_ptr<phx_batch> buildBatch(phx_ctx& c, size_t n, const int64_t* id, const double* epoch, const double* state6,
                                      const double* mass, const double* drag, const double* srp, const double* cd,
                                      const double* cr, double duration, const double* durations) {
    validateTargets(n, id, epoch, state6, mass, drag, srp, cd, cr);
    if (!durations && (!(duration > 0) || !std::isfinite(duration)))
        throw std::invalid_argument("duration_seconds must be finite and positive.");
    const double step = c.run.stepSeconds;
    auto b = std::make_unique<phx_batch>();
    b->ctx = &c;
    b->n = n;
    // FixedStep.h:7-29 — ET conversion, first step, launch count
    std::vector<double> start(n), endEt(n);
    long long count = 0;
    for (size_t i = 0; i < n; ++i) {
        const double d = durations ? durations[i] : duration;
        if (!(d > 0) || !std::isfinite(d)) throw std::invalid_argument("duration_seconds must be finite and positive.");
        start[i] = epoch[i] * 86400.0 + -43200.0;
        endEt[i] = (epoch[i] + d / 86400.0) * 86400.0 + -43200.0;
        const double span = endEt[i] - start[i];
        if (!(span > 0) || !std::isfinite(span) || std::ceil(span / step) >= 2147483647.0)
            throw std::invalid_argument("Duration/step cannot be represented by the integration time/index types.");
        count = std::max(count, (long long)std::ceil(span / step) + 1);
    }
    if (needsCoverageCheck(c)) {
        double lo = start[0], hi = endEt[0];
        for (size_t i = 0; i < n; ++i) {
            lo = std::min(lo, start[i]);
            hi = std::max(hi, endEt[i]);
        }
        checkEphemerisCoverage(c, lo, hi);
    }
    const long long launches = std::min(count, c.run.maxNumSteps);
    // epoch groups: identical (start, end) share one step sequence
    std::map<std::pair<double, double>, int> groupOf;
    std::vector<int> grp(n);
    for (size_t i = 0; i < n; ++i) {
        auto key = std::make_pair(start[i], endEt[i]);
        auto it = groupOf.find(key);
        if (it == groupOf.end()) it = groupOf.emplace(key, int(groupOf.size())).first;
        grp[i] = it->second;
    }
    const int G = int(groupOf.size());
    const bool grid = c.forceGrid || G > c.exactGroupsMax;
    b->grid = grid;
    std::vector<double> epochs;
    std::vector<long long> groupRecord(static_cast<size_t>(G), 0);
    int totalSteps = 0;
    double gridT0 = 0.0;
    int gridRows = 0;
    if (!grid) {
        std::vector<std::pair<double, double>> gkey(static_cast<size_t>(G));
        for (auto& kv : groupOf) gkey[size_t(kv.second)] = kv.first;
        // per-group step sequences; record epochs are the stage epochs themselves
        for (int g = 0; g < G; ++g) {
            groupRecord[size_t(g)] = (long long)epochs.size();  // record index of (step 0, stage 0)
            double cur = gkey[size_t(g)].first;
            const double end = gkey[size_t(g)].second;
            double dt = std::min(step, end - cur);
            int k = 0;
            while (k < launches) {
                for (int s = 0; s < kStages; ++s) {
                    double ts = cur;
                    if (s >= 1) ts += dt * kRkC[s - 1];
                    epochs.push_back(ts);
                }
                cur += dt;
                const double rem = end - cur;
                const bool fin = rem <= 0;
                dt = fin ? 0.0 : std::min(step, rem);
                ++k;
                if (fin) break;
            }
            totalSteps = std::max(totalSteps, k);
        }
    } else {
        // uniform grid over the whole batch window, two rows of margin each side
        double lo = start[0], hi = endEt[0];
        for (size_t i = 0; i < n; ++i) {
            lo = std::min(lo, start[i]);
            hi = std::max(hi, endEt[i]);
        }
        const double h = c.gridStep;
        gridT0 = lo - 2.0 * h;
        b->gridInvH = 1.0 / h;
        gridRows = int(std::ceil((hi - gridT0) / h)) + 3;
        for (int k = 0; k < gridRows; ++k) epochs.push_back(gridT0 + k * h);
        totalSteps = int(launches);
    }
    b->totalSteps = totalSteps;
    // order: by group (grid: by start epoch), then by altitude so drag lanes share tiles
    std::vector<int> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = int(i);
    std::vector<double> alt(n);
    for (size_t i = 0; i < n; ++i)
        alt[i] = std::sqrt(state6[i] * state6[i] + state6[n + i] * state6[n + i] + state6[2 * n + i] * state6[2 * n + i]);
    if (grid) {
        // longest work first, in classes of step counts 10 % wide (a tile's lanes end
        // within ~10 % of each other and the tail is short); inside a class by
        // radius, so the objects below the drag cutoff share tiles
        std::vector<int> cls(n);
        for (size_t i = 0; i < n; ++i)
            cls[i] = int(std::floor(std::log(std::max(1.0, std::ceil((endEt[i] - start[i]) / step))) / std::log(1.1)));
        std::stable_sort(order.begin(), order.end(), [&](int a, int b2) {
            if (cls[size_t(a)] != cls[size_t(b2)]) return cls[size_t(a)] > cls[size_t(b2)];
            return alt[size_t(a)] < alt[size_t(b2)];
        });
        for (size_t i = 0; i < n; ++i) grp[i] = 0;
    } else if (c.sortAltitude) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b2) {
            if (grp[size_t(a)] != grp[size_t(b2)]) return grp[size_t(a)] < grp[size_t(b2)];
            return alt[size_t(a)] < alt[size_t(b2)];
        });
    } else {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return grp[size_t(a)] < grp[size_t(b2)]; });
    }
    // tiles never mix groups
    std::vector<int> slotObj;  // slot -> object (-1 padding)
    std::vector<int> tileGroup;
    for (size_t k = 0; k < n;) {
        const int g = grp[size_t(order[k])];
        size_t e = k;
        while (e < n && grp[size_t(order[e])] == g) ++e;
        for (size_t t = k; t < e; t += kTile) {
            tileGroup.push_back(g);
            for (int l = 0; l < kTile; ++l) slotObj.push_back(t + size_t(l) < e ? order[t + size_t(l)] : -1);
        }
        k = e;
    }
    const int nTiles = int(tileGroup.size());
    const int nPad = nTiles * kTile;
    b->nTiles = nTiles;
    b->nPad = nPad;
    b->slotOf.assign(n, -1);
    const size_t np = static_cast<size_t>(nPad);
    std::vector<double> y(6 * np), cur(np), end(np), nd(np), ams(np), bcd(np);
    std::vector<uint8_t> st(np);
    std::vector<long long> tileRec(static_cast<size_t>(nTiles));
    for (int t = 0; t < nTiles; ++t) tileRec[size_t(t)] = groupRecord[size_t(tileGroup[size_t(t)])];
    // steps each tile needs at most (ceil(span / step) + 1 bounds the shortened-step sequence)
    b->tileSteps.assign(static_cast<size_t>(nTiles), 0);
    for (int sIdx = 0; sIdx < nPad; ++sIdx) {
        const int i = slotObj[size_t(sIdx)];
        if (i < 0) continue;
        const double span = endEt[size_t(i)] - start[size_t(i)];
        const long long need = std::min<long long>((long long)std::ceil(span / step) + 1, launches);
        int& ts = b->tileSteps[size_t(sIdx / kTile)];
        ts = std::max(ts, int(need));
    }
    for (int s = 0; s < nPad; ++s) {
        int i = slotObj[size_t(s)];
        const bool pad = i < 0;
        if (pad) i = slotObj[size_t(s - s % kTile)];  // copy of the tile's first object
        else b->slotOf[size_t(i)] = s;
        for (int d = 0; d < 6; ++d) y[size_t(d) * nPad + s] = state6[size_t(d) * n + i];
        cur[size_t(s)] = start[size_t(i)];
        end[size_t(s)] = endEt[size_t(i)];
        nd[size_t(s)] = std::min(step, endEt[size_t(i)] - start[size_t(i)]);
        // driver.cpp:117-120: area m^2 -> km^2, cr_eff = cr * srp / drag
        const double area = drag[i] * 1e-6;
        const double crEff = cr[i] * srp[i] / drag[i];
        ams[size_t(s)] = crEff * area / mass[i];
        bcd[size_t(s)] = cd[i] * area / mass[i];
        st[size_t(s)] = pad ? 4 : 0;
    }
    // device budget guard (Environment.h:977-1050): records + state must fit in 60 % of free memory
    size_t freeB = 0, totalB = 0;
    check(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    const size_t need = epochs.size() * (REC_SIZE + 1) * sizeof(double) + size_t(nPad) * (12 * sizeof(double) + 1);
    if (double(need) > 0.6 * double(freeB))
        throw std::length_error("Device memory budget exceeded (" + std::to_string(need >> 20) + " MiB needed); split into multiple simulations.");
    hipStream_t s = c.stream;
    b->y.upload(y.data(), y.size(), s);
    b->cur.upload(cur.data(), cur.size(), s);
    b->end.upload(end.data(), end.size(), s);
    b->nextDt.upload(nd.data(), nd.size(), s);
    b->ams.upload(ams.data(), ams.size(), s);
    b->bcd.upload(bcd.data(), bcd.size(), s);
    b->status.upload(st.data(), st.size(), s);
    b->tileRecord.upload(tileRec.data(), tileRec.size(), s);
    if (grid) {
        // one grid segment (record row 0) shared by every slot
        std::vector<long long> sr(np, 0);
        std::vector<double> st0v(np, gridT0);
        std::vector<int> rows(np, gridRows);
        b->slotRecord.upload(sr.data(), np, s);
        b->slotT0.upload(st0v.data(), np, s);
        b->slotRows.upload(rows.data(), np, s);
    }
    long long zero[4] = {0, 0, 0, 0};
    b->counters.upload(zero, 4, s);
    b->y0 = y;
    b->cur0 = cur;
    b->dt0 = nd;
    b->st0 = st;
    // epoch records
    DBuf<double> dEpochs;
    dEpochs.upload(epochs.data(), epochs.size(), s);
    b->records.alloc(epochs.size() * REC_SIZE);
    EpochArgs ea = epochArgs(c);
    ea.count = int(epochs.size());
    ea.epochs = dEpochs.p;
    ea.out = b->records.p;
    ea.gridRows = grid ? 1 : 0;
    check(hipEventRecord(c.ev0, s), "hipEventRecord");
    if (launchEpochRecords(ea, c.fm, s) != 0) throw HipError("epoch-record kernel launch failed");
    check(hipEventRecord(c.ev1, s), "hipEventRecord");
    check(hipStreamSynchronize(s), "epoch records");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, c.ev0, c.ev1), "hipEventElapsedTime");
    b->epochSeconds = ms * 1e-3;
    b->epochRecords = epochs.size();
    return b;
}

StepArgs contextStepArgs(const phx_ctx& c, int& W) {
    StepArgs a{};
    a.step = c.run.stepSeconds;
    a.shTerms = c.shTerms.p;
    a.shTerms32 = c.shTerms32.p;
    a.shOrderStart = c.shOrderStart.p;
    a.shDiag = c.shDiag.p;
    a.fp64Degree = c.run.precision.mixed ? c.run.precision.fp64Degree : -1;
    a.fp64Order = c.run.precision.fp64Order;
    a.nrlCoeff = c.nrl.p;
    a.weather = c.weather.p;
    a.wDays = c.env.hasWeather ? c.env.weather.days : 0;
    a.wStart = c.env.hasWeather ? c.env.weather.startMjdUtc : 0.0;
    a.nrlSlots = c.nrlSlots.p;
    a.gridInvH = 1.0 / c.gridStep;
    a.leaps = c.leaps.p;
    a.leapCount = c.leapCount;
    // drag models run at most four waves per tile (the NRLMSISE-00 path needs
    // the register budget of one wave per SIMD); without a field the force
    // families spread over at most four waves
    W = c.fm.hasSh ? (c.fm.drag ? std::min(c.nWaves, 4) : c.nWaves) : std::min(c.nWaves, 4);
    if (c.fm.hasSh) orderBlocks(c, W, a.blockStart, false);
    else
        for (int w = 0; w <= kMaxWaves; ++w) a.blockStart[w] = 0;
    return a;
}

int stepsToEnd(double startEt, double endEt, double step) {
    double cur = startEt;
    double dt = std::min(step, endEt - cur);
    int k = 0;
    for (;;) {
        cur += dt;
        ++k;
        const double rem = endEt - cur;
        if (rem <= 0) return k;
        dt = std::min(step, rem);
    }
}

int runBatch(phx_ctx& c, phx_batch& b, int maxSteps, phx_stats* stats) {
    const auto w0 = std::chrono::steady_clock::now();
    int W = 1;
    StepArgs a = contextStepArgs(c, W);
    a.nTiles = b.nTiles;
    a.nPad = b.nPad;
    a.y = b.y.p;
    a.cur = b.cur.p;
    a.end = b.end.p;
    a.nextDt = b.nextDt.p;
    a.status = b.status.p;
    a.ams = b.ams.p;
    a.bcd = b.bcd.p;
    a.records = b.records.p;
    a.tileRecord = b.tileRecord.p;
    a.counters = b.counters.p;
    a.grid = b.grid ? 1 : 0;
    a.gridInvH = b.gridInvH;
    a.slotRecord = b.slotRecord.p;
    a.slotT0 = b.slotT0.p;
    a.slotRows = b.slotRows.p;
    int remaining = b.totalSteps - b.stepsDone;
    if (maxSteps > 0) remaining = std::min(remaining, maxSteps);
    const long long replays0 = b.graphReplays, captures0 = b.graphCaptures;
    long long before = 0;
    check(hipMemcpyAsync(&before, b.counters.p, sizeof before, hipMemcpyDeviceToHost, c.stream), "counter read");
    check(hipEventRecord(c.ev0, c.stream), "hipEventRecord");
    int launches = 0;
    int kernelsPerStep = 0;
    int blocksFor = W;   // wave count the order blocks in `a` were split for
    int groupsMin = 1 << 30, groupsMax = 0;   // SH workgroups per tile (split path)
    // split stages: shared-epoch batches with drag, and per-lane epochs with the
    // eight-wave drag kernel
    const bool split = c.split && c.fm.drag && c.fm.hasSh && (!b.grid || c.splitSpecies == 8);
    if (split) {
        const size_t np = size_t(b.nPad);
        if (!b.ks.p) {
            b.ks.alloc(size_t(kStages) * 6 * np);
            b.ysBuf.alloc(7 * np);
            b.parts.alloc(6 * np);
            b.spec.alloc(8 * np);
            b.lane.alloc(8 * np);
            b.trig.alloc(6 * np);
        }
        const int Ws = c.splitWaves;
        orderBlocks(c, Ws, a.blockStart, true);
        SplitArgs sp{b.ks.p, b.ysBuf.p, b.parts.p, b.spec.p, b.lane.p, 0, 0, c.splitSpecies};
        sp.trig = b.trig.p;
        // the step index of exact records lives on the device (advanced by the
        // combine kernel), so every step launches the same kernels and arguments
        if (!b.stepDev.p) b.stepDev.alloc(1);
        // graphs captured under other options have expired
        for (size_t i = b.graphs.size(); i-- > 0;)
            if (b.graphs[i].gen != c.optGen) b.dropGraph(i);
        b.stepHost = b.stepsDone;
        check(hipMemcpyAsync(b.stepDev.p, &b.stepHost, sizeof(int), hipMemcpyHostToDevice, c.stream), "step index");
        sp.stepDev = b.stepDev.p;
        int groupsFor = 1;   // group count the group blocks in `sp` were split for
        // per stage: SH kernel, then either the drag kernel or species + finish kernels
        kernelsPerStep = kStages * (c.splitSpecies == 0 || c.splitSpecies == 8 ? 2 : 3) + 1;
        const int K = c.graphSteps;
        while (remaining > 0) {
            int active = b.nTiles;
            while (active > 0 && b.tileSteps[size_t(active - 1)] <= b.stepsDone) --active;
            if (active == 0) {
                b.stepsDone += remaining;
                break;
            }
            a.nTiles = active;
            sp.step = b.stepsDone;
            // while the active tiles leave CUs idle, share each tile's SH over
            // several workgroups
            int G = 1;
            if (Ws >= 8 && c.splitGroups > 1) G = std::max(1, std::min(c.splitGroups, c.numCU / std::max(1, active)));
            groupsMin = std::min(groupsMin, G);
            groupsMax = std::max(groupsMax, G);
            if (G > 1) {
                if (!b.tileCount.p) {
                    b.shPart.alloc(size_t(kMaxGroups) * 3 * np);
                    b.famPart.alloc(6 * np);
                    b.tileCount.alloc(size_t(b.nTiles));
                    check(hipMemsetAsync(b.tileCount.p, 0, size_t(b.nTiles) * sizeof(unsigned), c.stream), "hipMemsetAsync");
                }
                if (G != groupsFor) {
                    splitOrderBlocks(c, Ws, G, sp.groupBlockStart);
                    groupsFor = G;
                }
            }
            sp.groups = G;
            sp.shPart = b.shPart.p;
            sp.famPart = b.famPart.p;
            sp.tileCount = b.tileCount.p;
            // K steps as one graph while the active tiles stay the same for all of
            // them (the next tile ends no sooner than K steps from here); while the
            // whole batch is active, shorter runs replay graphs of a power of two
            // steps, captured together with the K-step graph, so that a call of any
            // length after the first reuses graphs instead of capturing
            const int untilChange = b.tileSteps[size_t(active - 1)] - b.stepsDone;
            const int room = std::min(remaining, untilChange);
            int k = 0;
            if (K > 0 && room >= K) k = K;
            else if (K > 0 && active == b.nTiles)
                for (int p = 2; p < K && p <= room; p *= 2) k = p;
            if (k > 0) {
                auto find = [&](int steps) -> int {
                    for (size_t i = 0; i < b.graphs.size(); ++i)
                        if (b.graphs[i].active == active && b.graphs[i].groups == G && b.graphs[i].steps == steps) return int(i);
                    return -1;
                };
                auto capture = [&](int steps) {
                    if (b.graphs.size() >= phx_batch::kMaxStepGraphs) b.dropGraph(0);
                    check(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                    int rc = 0;
                    for (int j = 0; j < steps && rc == 0; ++j) {
                        for (int s = 0; s < kStages && rc == 0; ++s) {
                            sp.stage = s;
                            rc = launchSplitStage(a, c.fm, sp, Ws, c.stream);
                        }
                        if (rc == 0) rc = launchSplitCombine(a, sp, c.stream);
                    }
                    hipGraph_t g = nullptr;
                    const hipError_t e = hipStreamEndCapture(c.stream, &g);
                    if (rc != 0 || e != hipSuccess) {
                        if (g) (void)hipGraphDestroy(g);
                        throw HipError(std::string("split step graph capture failed: ") + hipGetErrorString(e));
                    }
                    hipGraphExec_t x = nullptr;
                    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
                    if (ei != hipSuccess) {
                        (void)hipGraphDestroy(g);
                        throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
                    }
                    b.graphs.push_back({active, G, steps, c.optGen, g, x});
                    // the graph's device-side setup now, not on its first replay
                    check(hipGraphUpload(x, c.stream), "hipGraphUpload");
                    ++b.graphCaptures;
                };
                if (find(k) < 0) {
                    capture(k);
                    if (active == b.nTiles) {
                        if (k != K && find(K) < 0) capture(K);
                        for (int p = 2; p < K; p *= 2)
                            if (find(p) < 0) capture(p);
                    }
                }
                const int gi = find(k);
                if (gi < 0) throw HipError("split step graph missing after capture");
                check(hipGraphLaunch(b.graphs[size_t(gi)].exec, c.stream), "hipGraphLaunch");
                ++b.graphReplays;
                launches += k * kernelsPerStep;
                b.stepsDone += k;
                remaining -= k;
                continue;
            }
            for (int s = 0; s < kStages; ++s) {
                sp.stage = s;
                if (launchSplitStage(a, c.fm, sp, Ws, c.stream) != 0) throw HipError("split stage launch failed");
            }
            if (launchSplitCombine(a, sp, c.stream) != 0) throw HipError("split combine launch failed");
            launches += kernelsPerStep;
            ++b.stepsDone;
            --remaining;
        }
        W = Ws;
    }
    while (remaining > 0) {
        const int k = std::min(remaining, std::max(1, c.stepsPerLaunch));
        // launch only the prefix of tiles that still has work (uneven batches are
        // ordered longest work first); when those tiles cannot give every SIMD a
        // wave, split each tile over at least four waves
        int active = b.nTiles;
        while (active > 0 && b.tileSteps[size_t(active - 1)] <= b.stepsDone) --active;
        if (active == 0) {
            b.stepsDone += remaining;
            break;
        }
        int Wl = W;
        if (c.fm.hasSh && W < 4 && (long long)active * W < 4LL * c.numCU) Wl = 4;
        if (c.fm.hasSh && Wl != blocksFor) {
            orderBlocks(c, Wl, a.blockStart, false);
            blocksFor = Wl;
        }
        a.nTiles = active;
        a.stepBase = b.stepsDone;
        a.steps = k;
        const int rc = launchStep(a, c.fm, Wl, c.stream);
        if (rc != 0) throw HipError("step kernel launch failed");
        b.stepsDone += k;
        remaining -= k;
        ++launches;
    }
    check(hipEventRecord(c.ev1, c.stream), "hipEventRecord");
    long long after = 0;
    check(hipMemcpyAsync(&after, b.counters.p, sizeof after, hipMemcpyDeviceToHost, c.stream), "counter read");
    check(hipStreamSynchronize(c.stream), "step kernels");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, c.ev0, c.ev1), "hipEventElapsedTime");
    if (stats) {
        stats->object_steps = double(after - before);
        stats->kernel_seconds = ms * 1e-3;
        stats->epoch_seconds = b.epochSeconds;
        stats->epoch_records = double(b.epochRecords);
        stats->launches = launches;
        stats->waves = W;
        stats->steps_per_launch = split ? 1 : c.stepsPerLaunch;
        stats->kernels_per_step = kernelsPerStep;
        stats->sh_groups_min = groupsMax ? groupsMin : 0;
        stats->sh_groups_max = groupsMax;
        stats->graph_replays = double(b.graphReplays - replays0);
        stats->graph_captures = double(b.graphCaptures - captures0);
        stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    }
    return 0;
}

void downloadBatch(phx_ctx& c, phx_batch& b, double* outEpoch, double* outState, bool requireEnd = true) {
    const size_t np = static_cast<size_t>(b.nPad);
    std::vector<double> y(6 * np), cur(np);
    std::vector<uint8_t> st(np);
    check(hipMemcpyAsync(y.data(), b.y.p, y.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream), "D2H");
    check(hipMemcpyAsync(cur.data(), b.cur.p, cur.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream), "D2H");
    check(hipMemcpyAsync(st.data(), b.status.p, st.size(), hipMemcpyDeviceToHost, c.stream), "D2H");
    check(hipStreamSynchronize(c.stream), "download");
    for (size_t i = 0; i < b.n; ++i) {
        const int s = b.slotOf[i];
        if (requireEnd && !(st[size_t(s)] & 2))
            throw std::range_error("A target stopped without reaching the requested end of simulation.");
        if (outEpoch) outEpoch[i] = cur[size_t(s)] * (1.0 / 86400.0) + 0.5;   // TimeConversion.h:36-40
        if (outState)
            for (int d = 0; d < 6; ++d) outState[size_t(d) * b.n + i] = y[size_t(d) * b.nPad + s];
    }
}

}  // namespace phx

extern "C" {

const char* phx_version(void) { return "phx 0.1 (gfx950)"; }

int phx_create(const char* config_json_path, phx_ctx** out) {
    if (!out || !config_json_path) return PHX_EINVAL;
    *out = nullptr;
    auto c = std::make_unique<phx_ctx>();
    try {
        const std::string path = config_json_path;
        json::Value root = json::readFile(path);
        parseConfig(root, dirName(path), c->run, c->env.model, false);
        loadEnvironment(c->env);
        buildForceModel(*c);  // host-only; rejects what the kernels cannot hold before any device call
        int ndev = 0;
        const hipError_t de = hipGetDeviceCount(&ndev);
        if (de != hipSuccess || ndev <= 0)
            throw HipError(std::string("No usable HIP device: ") + hipGetErrorString(de));
        const char* env = std::getenv("PHX_DATA_DIR");
        c->dataDir = env ? env : libraryDir() + "/data";
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        c->numCU = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&c->ev0), "hipEventCreate");
        check(hipEventCreate(&c->ev1), "hipEventCreate");
        if (uploadTideFactors() != 0) throw HipError("tide factor upload failed");
        if (noallocInit() != 0) throw HipError("device scratch array of phx_propagate_noalloc not found");
        loadTables(*c);
        computeLeaps(*c);
        if (c->env.hasSh) buildShTables(*c);
        buildRotTables(*c);
        if (!c->env.brie.empty()) {
            std::vector<double> rows(size_t(kEphSlots) * kEphRow, 0.0);
            const ModelConfig& m = c->env.model;
            for (int slot = 0; slot < kEphSlots; ++slot) {
                const int naif = ephSlotBody(m, slot);
                if (naif < 0) continue;
                const auto ch = brieChain(c->env.brie, naif, NAIF_EARTH);
                double* row = rows.data() + size_t(slot) * kEphRow;
                row[0] = double(ch.size());
                for (size_t k = 0; k < ch.size(); ++k) {
                    const BrieUnit& u = c->env.brie.units[size_t(ch[k].unit)];
                    row[1 + 4 * k] = double(u.offset);
                    row[2 + 4 * k] = double(u.nIntervals);
                    row[3 + 4 * k] = double(u.pdeg);
                    row[4 + 4 * k] = ch[k].sign;
                }
            }
            c->eph.upload(c->env.brie.data.data(), c->env.brie.data.size(), c->stream);
            c->ephChain.upload(rows.data(), rows.size(), c->stream);
        }
        if (!c->env.brot.empty()) {
            const BrotIers& u = c->env.brot.iers[0];
            const size_t len = size_t(3 * u.nNut + 5 * u.nErp + 4 + u.nErp);
            c->iers.upload(c->env.brot.data.data() + u.offset, len, c->stream);
        }
        if (c->fm.drag && c->env.hasWeather) {
            // weather-slot rows for the uneven-epoch path
            const int rows = c->env.weather.days * 8;
            c->nrlSlots.alloc(size_t(rows) * kNrlSlotSize);
            if (launchNrlSlots(epochArgs(*c), c->nrlSlots.p, rows, c->stream) != 0) throw HipError("weather-slot kernel launch failed");
        }
        check(hipStreamSynchronize(c->stream), "table upload");
        *out = c.release();
        return PHX_OK;
    } catch (const std::exception& e) {
        g_createError = e.what();
        return dynamic_cast<const HipError*>(&e) ? PHX_EDEVICE : PHX_EINVAL;
    }
}

const char* phx_last_error(const phx_ctx* ctx) {
    // a copy per calling thread: another thread's serving call may replace the
    // context's message meanwhile
    thread_local std::string copy;
    copy = ctx ? ctx->lastError() : g_createError;
    return copy.c_str();
}

void phx_destroy(phx_ctx* ctx) {
    if (!ctx) return;
    ctx->server.reset();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->shTerms.release();
    ctx->shTerms32.release();
    ctx->shOrderStart.release();
    ctx->shDiag.release();
    ctx->nrl.release();
    ctx->freq.release();
    ctx->waves.release();
    ctx->terms.release();
    ctx->eop.release();
    ctx->weather.release();
    ctx->nrlSlots.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int phx_set_option(phx_ctx* c, const char* key, double v) {
    if (!c || !key) return PHX_EINVAL;
    const std::string k = key;
    if (k == "waves") {
        const int w = int(v);
        if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) { c->setError("waves must be 1, 2, 4, 8 or 16"); return PHX_EINVAL; }
        c->nWaves = w;
    } else if (k == "steps_per_launch") {
        if (!(v >= 1)) { c->setError("steps_per_launch must be >= 1"); return PHX_EINVAL; }
        c->stepsPerLaunch = int(v);
    } else if (k == "share0") c->share0 = v;
    else if (k == "share_corr") c->shareCorr = v;
    else if (k == "sort_altitude") c->sortAltitude = v != 0.0;
    else if (k == "exact_groups_max") c->exactGroupsMax = int(v);
    else if (k == "grid_step") {
        if (!(v > 0 && v <= 600)) { c->setError("grid_step must be in (0, 600] s"); return PHX_EINVAL; }
        c->gridStep = v;
    } else if (k == "force_grid") c->forceGrid = v != 0.0;
    else if (k == "split") c->split = v != 0.0;
    else if (k == "split_groups") {
        if (!(v >= 1.0 && v <= double(kMaxGroups))) { c->setError("split_groups must be 1 .. 4"); return PHX_EINVAL; }
        c->splitGroups = int(v);
    }
    else if (k == "graph_steps") {
        if (!(v >= 0.0 && v <= 256.0)) { c->setError("graph_steps must be 0 .. 256"); return PHX_EINVAL; }
        c->graphSteps = int(v);
    }
    else if (k == "split_species") {
        if (v != 0 && v != 1 && v != 2 && v != 4 && v != 8) { c->setError("split_species must be 0, 1, 2, 4 or 8"); return PHX_EINVAL; }
        c->splitSpecies = int(v);
    }
    else if (k == "split_cost_pg" || k == "split_cost_corr" || k == "split_cost_srp" || k == "split_cost_lane" ||
             k == "split_cost_order" || k == "split_cost_order0") {
        if (!(v >= 0.0 && v <= 1e4)) { c->setError(k + " must be in [0, 10000]"); return PHX_EINVAL; }
        (k == "split_cost_pg"      ? c->splitCostPg
         : k == "split_cost_corr"  ? c->splitCostCorr
         : k == "split_cost_srp"   ? c->splitCostSrp
         : k == "split_cost_lane"  ? c->splitCostLane
         : k == "split_cost_order" ? c->splitCostOrder
                                   : c->splitCostOrder0) = v;
    } else if (k == "split_waves") {
        if (v != 4 && v != 8 && v != 16) { c->setError("split_waves must be 4, 8 or 16"); return PHX_EINVAL; }
        c->splitWaves = int(v);
    }
    else if (k == "serve_steps_per_tick") {
        if (!(v >= 1)) { c->setError("serve_steps_per_tick must be >= 1"); return PHX_EINVAL; }
        c->serveStepsPerTick = int(v);
    } else if (k == "serve_slots" || k == "serve_record_mib") {
        std::lock_guard<std::mutex> lk(c->serverMu);
        if (c->server) { c->setError(k + " must be set before the first phx_submit"); return PHX_EINVAL; }
        if (!(v >= 1)) { c->setError(k + " must be >= 1"); return PHX_EINVAL; }
        if (k == "serve_slots") c->serveSlots = (long long)v;
        else c->serveRecordMiB = v;
    } else { c->setError("unknown option " + k); return PHX_EINVAL; }
    ++c->optGen;
    return PHX_OK;
}

double phx_get_option(const phx_ctx* c, const char* key) {
    if (!c || !key) return std::nan("");
    const std::string k = key;
    if (k == "waves") return c->nWaves;
    if (k == "steps_per_launch") return c->stepsPerLaunch;
    if (k == "share0") return c->share0;
    if (k == "share_corr") return c->shareCorr;
    if (k == "sort_altitude") return c->sortAltitude;
    if (k == "exact_groups_max") return c->exactGroupsMax;
    if (k == "grid_step") return c->gridStep;
    if (k == "force_grid") return c->forceGrid;
    if (k == "split_groups") return c->splitGroups;
    if (k == "graph_steps") return c->graphSteps;
    if (k == "split") return c->split;
    if (k == "split_species") return c->splitSpecies;
    if (k == "split_waves") return c->splitWaves;
    if (k == "split_cost_pg") return c->splitCostPg;
    if (k == "split_cost_corr") return c->splitCostCorr;
    if (k == "split_cost_srp") return c->splitCostSrp;
    if (k == "split_cost_lane") return c->splitCostLane;
    if (k == "split_cost_order") return c->splitCostOrder;
    if (k == "split_cost_order0") return c->splitCostOrder0;
    if (k == "step_seconds") return c->run.stepSeconds;
    if (k == "duration_seconds") return c->run.durationSeconds;
    if (k == "max_num_steps") return double(c->run.maxNumSteps);
    if (k == "sh_degree") return c->fm.hasSh ? c->fm.shDegree : 0;
    if (k == "mixed") return c->run.precision.mixed ? 1 : 0;
    if (k == "serve_steps_per_tick") return c->serveStepsPerTick;
    if (k == "serve_record_mib") return c->serveRecordMiB;
    if (k.rfind("serve_", 0) == 0) {
        std::lock_guard<std::mutex> lk(c->serverMu);
        if (k == "serve_slots" && !c->server) return double(c->serveSlots);
        return c->server ? c->server->stat(k) : 0.0;
    }
    return std::nan("");
}

int phx_batch_create(phx_ctx* c, size_t n, const int64_t* id, const double* epoch, const double* state6, const double* mass,
                     const double* drag, const double* srp, const double* cd, const double* cr, double duration,
                     phx_batch** out) {
    if (!c || !out) return PHX_EINVAL;
    try {
        *out = buildBatch(*c, n, id, epoch, state6, mass, drag, srp, cd, cr, duration, nullptr).release();
        return PHX_OK;
    } catch (const std::length_error& e) {
        c->setError(e.what());
        return PHX_ENOMEM;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

int phx_batch_create_uneven(phx_ctx* c, size_t n, const int64_t* id, const double* epoch, const double* state6,
                            const double* mass, const double* drag, const double* srp, const double* cd, const double* cr,
                            const double* duration_s, phx_batch** out) {
    if (!c || !out || !duration_s) return PHX_EINVAL;
    try {
        *out = buildBatch(*c, n, id, epoch, state6, mass, drag, srp, cd, cr, 0.0, duration_s).release();
        return PHX_OK;
    } catch (const std::length_error& e) {
        c->setError(e.what());
        return PHX_ENOMEM;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

int phx_batch_run(phx_batch* b, int maxSteps, phx_stats* stats) {
    if (!b) return PHX_EINVAL;
    try {
        return runBatch(*b->ctx, *b, maxSteps, stats);
    } catch (const std::exception& e) {
        return fail(b->ctx, e);
    }
}

int phx_batch_autotune(phx_batch* b, int trialSteps, phx_stats* best) {
    if (!b || trialSteps < 1) return PHX_EINVAL;
    phx_ctx& c = *b->ctx;
    try {
        struct Cand {
            int waves;
            double share0;
            int split;    // split-stage path (shared-epoch batches with drag)
            int splitWaves;
            int splitSpecies;
        };
        std::vector<Cand> cands;
        if (c.fm.hasSh) {
            for (int w : {1, 2, 4, 8, 16})
                for (double s0 : {0.7, 0.85, 1.0}) {
                    if (w == 1 && s0 != 1.0) continue;
                    if (c.fm.drag && w > 4) continue;
                    cands.push_back({w, s0, 0, c.splitWaves, c.splitSpecies});
                }
            if (c.fm.drag)
                for (int sw : {8, 16})
                    for (int ss : {1, 8})
                        if (!b->grid || ss == 8) cands.push_back({c.nWaves, c.share0, 1, sw, ss});
        } else {
            for (int w : {1, 2, 4}) cands.push_back({w, 1.0, 0, c.splitWaves, c.splitSpecies});
        }
        const Cand keep{c.nWaves, c.share0, c.split ? 1 : 0, c.splitWaves, c.splitSpecies};
        auto apply = [&c](const Cand& k) {
            ++c.optGen;
            c.nWaves = k.waves;
            c.share0 = k.share0;
            c.split = k.split != 0;
            c.splitWaves = k.splitWaves;
            c.splitSpecies = k.splitSpecies;
        };
        double bestT = 0.0;
        int bestI = -1;
        phx_stats st{}, bestSt{};
        for (size_t i = 0; i < cands.size(); ++i) {
            apply(cands[i]);
            int rc = phx_batch_reset(b);
            if (rc) throw std::runtime_error(c.lastError());
            runBatch(c, *b, trialSteps, &st);
            if (st.object_steps <= 0) continue;
            const double t = st.kernel_seconds / st.object_steps;
            if (bestI < 0 || t < bestT) {
                bestT = t;
                bestI = int(i);
                bestSt = st;
            }
        }
        apply(bestI < 0 ? keep : cands[size_t(bestI)]);
        const int rc = phx_batch_reset(b);
        if (best) *best = bestSt;
        return rc;
    } catch (const std::exception& e) {
        return fail(&c, e);
    }
}

int phx_batch_done(const phx_batch* b) { return b && b->stepsDone >= b->totalSteps ? 1 : 0; }

int phx_batch_download(phx_batch* b, double* outEpoch, double* outState) {
    if (!b) return PHX_EINVAL;
    try {
        downloadBatch(*b->ctx, *b, outEpoch, outState);
        return PHX_OK;
    } catch (const std::range_error& e) {
        b->ctx->setError(e.what());
        return PHX_ENOTDONE;
    } catch (const std::exception& e) {
        return fail(b->ctx, e);
    }
}

int phx_batch_state(phx_batch* b, double* outEpoch, double* outState) {
    if (!b) return PHX_EINVAL;
    try {
        downloadBatch(*b->ctx, *b, outEpoch, outState, false);
        return PHX_OK;
    } catch (const std::exception& e) {
        return fail(b->ctx, e);
    }
}

int phx_batch_reset(phx_batch* b) {
    if (!b) return PHX_EINVAL;
    try {
        phx_ctx& c = *b->ctx;
        check(hipMemcpyAsync(b->y.p, b->y0.data(), b->y0.size() * sizeof(double), hipMemcpyHostToDevice, c.stream), "H2D");
        check(hipMemcpyAsync(b->cur.p, b->cur0.data(), b->cur0.size() * sizeof(double), hipMemcpyHostToDevice, c.stream), "H2D");
        check(hipMemcpyAsync(b->nextDt.p, b->dt0.data(), b->dt0.size() * sizeof(double), hipMemcpyHostToDevice, c.stream), "H2D");
        check(hipMemcpyAsync(b->status.p, b->st0.data(), b->st0.size(), hipMemcpyHostToDevice, c.stream), "H2D");
        check(hipStreamSynchronize(c.stream), "reset");
        b->stepsDone = 0;
        return PHX_OK;
    } catch (const std::exception& e) {
        return fail(b->ctx, e);
    }
}

void phx_batch_destroy(phx_batch* b) {
    if (!b) return;
    if (b->ctx && b->ctx->stream) (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int phx_propagate(phx_ctx* c, size_t n, const int64_t* id, const double* epoch, const double* state6, const double* mass,
                  const double* drag, const double* srp, const double* cd, const double* cr, double duration,
                  double* outEpoch, double* outState, phx_stats* stats) {
    if (!c) return PHX_EINVAL;
    phx_batch* b = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = phx_batch_create(c, n, id, epoch, state6, mass, drag, srp, cd, cr, duration, &b);
    if (rc != PHX_OK) return rc;
    rc = phx_batch_run(b, 0, stats);
    if (rc == PHX_OK) rc = phx_batch_download(b, outEpoch, outState);
    phx_batch_destroy(b);
    if (stats) stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int phx_submit(phx_ctx* c, const phx_request* req, uint64_t* ticket) {
    if (!c || !req || !ticket) return PHX_EINVAL;
    try {
        Server* s = nullptr;
        {
            std::lock_guard<std::mutex> lk(c->serverMu);
            if (!c->server) c->server = std::make_unique<Server>(*c);
            s = c->server.get();
        }
        *ticket = s->submit(*req);
        return PHX_OK;
    } catch (const std::length_error& e) {
        c->setError(e.what());
        return PHX_ENOMEM;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

int phx_poll(phx_ctx* c, uint64_t ticket, phx_result* result) {
    if (!c) return PHX_EINVAL;
    Server* s = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->serverMu);
        s = c->server.get();
    }
    if (!s) {
        c->setError("no request was submitted");
        return PHX_EINVAL;
    }
    std::string err;
    const int rc = s->poll(ticket, result, err);
    if (rc != PHX_OK && rc != PHX_EAGAIN) c->setError(err);
    return rc;
}

int phx_epoch_record(phx_ctx* c, double et, double* record, size_t len) {
    if (!c || !record || len < size_t(REC_SIZE)) return PHX_EINVAL;
    try {
        DBuf<double> e, out;
        e.upload(&et, 1, c->stream);
        out.alloc(REC_SIZE);
        EpochArgs a = epochArgs(*c);
        a.count = 1;
        a.epochs = e.p;
        a.out = out.p;
        if (launchEpochRecords(a, c->fm, c->stream) != 0) throw HipError("epoch-record kernel launch failed");
        check(hipMemcpyAsync(record, out.p, REC_SIZE * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "epoch record");
        return PHX_OK;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

int phx_force_probe(phx_ctx* c, size_t n, const double* pos, const double* vel, const double* mass, const double* area,
                    const double* crEff, const double* cd, double et, int which, double* acc) {
    if (!c || n == 0) return PHX_EINVAL;
    try {
        std::vector<double> ams(n), bcd(n), rec(REC_SIZE);
        for (size_t i = 0; i < n; ++i) {
            ams[i] = crEff[i] * area[i] / mass[i];
            bcd[i] = cd[i] * area[i] / mass[i];
        }
        int rc = phx_epoch_record(c, et, rec.data(), rec.size());
        if (rc) return rc;
        DBuf<double> dp, dv, da, db, dr, dacc;
        dp.upload(pos, 3 * n, c->stream);
        dv.upload(vel, 3 * n, c->stream);
        da.upload(ams.data(), n, c->stream);
        db.upload(bcd.data(), n, c->stream);
        dr.upload(rec.data(), rec.size(), c->stream);
        dacc.alloc(3 * n);
        if (launchForceProbe(int(n), dp.p, dv.p, da.p, db.p, dr.p, which, c->fm, c->shTerms.p, c->shOrderStart.p, c->shDiag.p,
                             c->nrl.p, dacc.p, c->stream) != 0)
            throw HipError("force probe launch failed");
        check(hipMemcpyAsync(acc, dacc.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "force probe");
        return PHX_OK;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

int phx_nrlmsise_probe(phx_ctx* c, size_t n, const double* in, int useAp, double* out) {
    if (!c || n == 0) return PHX_EINVAL;
    try {
        if (!c->nrl.p) {
            // load the coefficient table even when the configuration has no drag
            std::ifstream f(c->dataDir + "/nrlmsise00_coefficients.txt");
            if (!f) throw std::runtime_error("missing NRLMSISE-00 coefficient table");
            std::vector<double> v;
            std::string line;
            while (std::getline(f, line)) {
                if (line.empty() || line[0] == '#' || std::isalpha((unsigned char)line[0])) continue;
                std::istringstream s(line);
                double x;
                while (s >> x) v.push_back(x);
            }
            if (v.size() != 3300) throw std::runtime_error("NRLMSISE-00 coefficient table must hold 3300 values");
            appendNrlHot(v);
            c->nrl.upload(v.data(), v.size(), c->stream);
        }
        DBuf<double> di, dout;
        di.upload(in, 16 * n, c->stream);
        dout.alloc(11 * n);
        if (launchNrlmsiseProbe(int(n), di.p, useAp, c->nrl.p, dout.p, c->stream) != 0) throw HipError("probe launch failed");
        check(hipMemcpyAsync(out, dout.p, 11 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "nrlmsise probe");
        return PHX_OK;
    } catch (const std::exception& e) {
        return fail(c, e);
    }
}

}  // extern "C"
