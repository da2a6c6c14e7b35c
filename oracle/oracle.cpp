// ORACLE — test infrastructure only. CPU restatement of the reference host
// propagation path; the parity checker for the HIP library and the timed CPU
// baseline. Compiled with -ffp-contract=off so the operation order written
// here is the order executed.
//
// Restated reference files (src/common/engine/ unless noted):
//   paraHPOP/propagators/FixedStep.h:7-47          prepare / finishStep / postProcess
//   paraHPOP/propagators/host/Propagator.h:13-34   OpenMP loop over objects per step
//   third_party/parm/parm/integrate/rk/host/Step.h:33-124  stage recursion, final sum
//   paraHPOP/model/math/ode/RefODE.h:65-173        k = [v; a(t, r, v, params)]
//   paraHPOP/model/accelerations/PrivateAccelerations.h:29-209  force order
//   paraHPOP/model/accelerations/EarthCorrections.h:13-103      corrections wrapper
//   src/common/driver.cpp:99-123                   target validation and unit conversion
#include "oracle.h"

#include <omp.h>

#include <chrono>
#include <cstring>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../parahpop_amd/csrc/model.h"
#include "oracle_physics.h"
#include "oracle_nrlmsise.h"

using namespace oracle;

#ifdef ORACLE_FLOP_COUNT
thread_local oracle_flop::Counts oracle_flop::counts{};
thread_local int oracle_flop::category = oracle_flop::kPerObject;
#endif

struct orc_ctx {
    phx::RunConfig run;
    phx::Environment env;
    ShTables sh;
    bool hasSh = false;
    Eop eop;
    TideTables tides;
    TideConstants tideK;
    NrlmsiseTables nrlTables;
    NrlWeather weather;
    bool anyOcculter = false;
    std::map<int, std::vector<phx::BrieStep>> ephChains;   // Chebyshev ephemeris chains to the Earth
    std::vector<ShTables> rotSh;   // fields of the bodies other than the Earth (env.rot order)
};

// ICRF -> axes of the k-th body field other than the Earth's.
M3 rotBodyAxes(const orc_ctx& c, size_t k, double et) {
    ORACLE_FLOP_SCOPE(kEpoch);
    const phx::RotBody& r = c.env.rot[k];
    const phx::BrotBody& u = c.env.brot.bodies[size_t(r.unit)];
    const double* d = c.env.brot.data.data() + u.offset;
    if (u.type == 3) return binaryFrameAxes(d, u.p[2], u.p[3], u.p[0], et);
    const int deg[3] = {u.p[0], u.p[1], u.p[2]}, np[3] = {u.p[5], u.p[6], u.p[7]};
    const double* ang = nullptr;
    int angDeg = 0;
    if (r.bary >= 0) {
        const phx::BrotBody& b = c.env.brot.bodies[size_t(r.bary)];
        ang = c.env.brot.data.data() + b.offset;
        angDeg = b.p[0];
    }
    return iauBodyAxes(d, deg, np, ang, angDeg, et);
}

// ICRF -> ITRF rotation and Earth rate: the configured IERS2000 unit, or the
// analytic model (with the EOP table when loaded).
Orientation contextOrientation(const orc_ctx& c, double et) {
    ORACLE_FLOP_SCOPE(kEpoch);
    if (!c.env.brot.empty()) {
        const phx::BrotIers& u = c.env.brot.iers[0];
        return iers2000Orientation(c.env.brot.data.data() + u.offset, u.nNut, u.nErp, et);
    }
    return earthOrientation(c.eop.days >= 2 ? &c.eop : nullptr, et);
}

// Geocentric position (km, ICRF) of a body: the configured Chebyshev
// ephemeris (brie.h), or the analytic series.
V3 bodyPosition(const orc_ctx& c, int naif, double et) {
    ORACLE_FLOP_SCOPE(kEpoch);
    if (c.ephChains.empty()) return bodyGeocentric(naif, et);
    auto it = c.ephChains.find(naif);
    if (it == c.ephChains.end()) throw std::runtime_error("no ephemeris chain for body " + std::to_string(naif));
    double p[3];
    phx::brieChainPosition(c.env.brie, it->second, et, p);
    return V3{p[0], p[1], p[2]};
}

static thread_local std::string g_error;

const char* orc_last_error(void) { return g_error.c_str(); }

static std::string dirName(const std::string& p) {
    const size_t k = p.find_last_of('/');
    if (k == std::string::npos) return ".";
    return k == 0 ? "/" : p.substr(0, k);
}

static void loadTideTables(const std::string& dir, TideTables& t) {
    {
        std::ifstream in(dir + "/iers2010_solid_tide_frequency.txt");
        if (!in) throw std::runtime_error("missing tide frequency table in " + dir);
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream s(line);
            FreqTerm f;
            s >> f.order >> f.mult[0] >> f.mult[1] >> f.mult[2] >> f.mult[3] >> f.mult[4] >> f.inPhase >> f.outPhase;
            if (!s || f.order < 0 || f.order > 2) throw std::runtime_error("bad tide frequency row");
            t.freq[f.order].push_back(f);
        }
    }
    {
        std::ifstream in(dir + "/fes2004_ocean_tide.txt");
        if (!in) throw std::runtime_error("missing FES2004 table in " + dir);
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream s(line);
            char tag;
            s >> tag;
            if (tag == 'W') {
                OceanWave w;
                for (int& d : w.doodson) s >> d;
                s >> w.begin >> w.end;
                t.waves.push_back(w);
            } else if (tag == 'T') {
                OceanTerm o;
                s >> o.n >> o.m >> o.cp >> o.sp >> o.cm >> o.sm;
                t.terms.push_back(o);
            }
            if (!s) throw std::runtime_error("bad FES2004 row");
        }
    }
}

int orc_create(const char* config_json_path, const char* data_dir, orc_ctx** out) {
    try {
        auto ctx = std::make_unique<orc_ctx>();
        const std::string cfgPath = config_json_path;
        phx::json::Value root = phx::json::readFile(cfgPath);
        phx::parseConfig(root, dirName(cfgPath), ctx->run, ctx->env.model, false);
        phx::loadEnvironment(ctx->env);
        if (ctx->env.hasSh) {
            const phx::ShField& f = ctx->env.sh;
            ctx->sh = ShTables{f.degree, f.order, f.normalized, f.radius, f.gm, f.C.data(), f.S.data(), f.factors.data()};
            ctx->hasSh = true;
        }
        for (const auto& r : ctx->env.rot) {
            const phx::ShField& f = r.field;
            ctx->rotSh.push_back(ShTables{f.degree, f.order, f.normalized, f.radius, f.gm, f.C.data(), f.S.data(),
                                          f.factors.data()});
        }
        if (ctx->env.hasEop) {
            const phx::EopTable& e = ctx->env.eop;
            ctx->eop.first = e.firstMjdUtc;
            ctx->eop.days = e.days;
            ctx->eop.xp = e.xp;
            ctx->eop.yp = e.yp;
            ctx->eop.dut = e.dut;
            ctx->eop.dat = e.dat;
        }
        const std::string dir = data_dir ? data_dir : ".";
        if (ctx->env.model.tidesActive()) loadTideTables(dir, ctx->tides);
        ctx->tideK.zeroTide = ctx->env.model.zeroTide;
        if (ctx->env.hasWeather) {
            loadNrlmsiseTables(dir + "/nrlmsise00_coefficients.txt", ctx->nrlTables);
            const phx::WeatherTable& w = ctx->env.weather;
            ctx->weather.startMjdUtc = w.startMjdUtc;
            ctx->weather.days = w.days;
            ctx->weather.data = w.data;
        }
        for (const auto& b : ctx->env.model.bodies)
            if (b.occulting && b.naif != phx::NAIF_SUN) ctx->anyOcculter = true;
        if (!ctx->env.brie.empty()) {
            for (const auto& b : ctx->env.model.bodies)
                ctx->ephChains[b.naif] = phx::brieChain(ctx->env.brie, b.naif, phx::NAIF_EARTH);
            for (int naif : {phx::NAIF_SUN, phx::NAIF_MOON})
                if (!ctx->ephChains.count(naif) && ctx->env.model.tidesActive())
                    ctx->ephChains[naif] = phx::brieChain(ctx->env.brie, naif, phx::NAIF_EARTH);
        }
        *out = ctx.release();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

void orc_destroy(orc_ctx* ctx) { delete ctx; }
double orc_step_seconds(const orc_ctx* ctx) { return ctx->run.stepSeconds; }
double orc_duration_seconds(const orc_ctx* ctx) { return ctx->run.durationSeconds; }

namespace {

struct Params { double mass, area, cr, cd; };

// accelerations/EarthCorrections.h:72-103 (COI = Earth) and HpopCorrections.h:378-392
V3 earthCorrections(const orc_ctx& c, V3 pos, V3 vel, double et, const Orientation* ori) {
    const phx::ModelConfig& m = c.env.model;
    if (!m.anyCorrections()) return V3{};
    TideInput in{};
    in.r[0] = 1000.0 * pos.x; in.r[1] = 1000.0 * pos.y; in.r[2] = 1000.0 * pos.z;
    in.v[0] = 1000.0 * vel.x; in.v[1] = 1000.0 * vel.y; in.v[2] = 1000.0 * vel.z;
    if (m.tidesActive()) {
        const TimeData t = timeAt(c.eop, et);
        in.mjdUtc = t.mjdUtc; in.ut1Utc = t.ut1Utc; in.ttUtc = t.ttUtc;
        in.xpArcsec = t.xpArcsec; in.ypArcsec = t.ypArcsec;
        if (!t.valid) return V3{std::nan(""), std::nan(""), std::nan("")};
        const V3 sun = bodyPosition(c, phx::NAIF_SUN, et), moon = bodyPosition(c, phx::NAIF_MOON, et);
        in.sun[0] = 1000.0 * sun.x; in.sun[1] = 1000.0 * sun.y; in.sun[2] = 1000.0 * sun.z;
        in.moon[0] = 1000.0 * moon.x; in.moon[1] = 1000.0 * moon.y; in.moon[2] = 1000.0 * moon.z;
        for (int i = 0; i < 9; ++i) in.E[i] = ori->R.m[i];
    }
    double solid[3] = {0, 0, 0}, ocean[3] = {0, 0, 0}, rel[3] = {0, 0, 0};
    CS7 coeff;
    if (m.solidTides) {
        solidCoefficients(in, c.tideK, c.tides, coeff);
        tideAcceleration(in, c.tideK, coeff, solid, 4);
    }
    if (m.oceanTides) {
        oceanCoefficients(in, c.tides, coeff);
        tideAcceleration(in, c.tideK, coeff, ocean, 6);
    }
    if (m.relativity) relativity(in, c.tideK, rel);
    return V3{1.0e-3 * (solid[0] + ocean[0] + rel[0]), 1.0e-3 * (solid[1] + ocean[1] + rel[1]),
              1.0e-3 * (solid[2] + ocean[2] + rel[2])};
}

// accelerations/PrivateAccelerations.h:29-209 with COI = Earth.
// `only` >= 0 restricts the sum to one force family (unit tests).
V3 acceleration(const orc_ctx& c, V3 pos, V3 vel, double et, const Params& p, int only = -1) {
    const phx::ModelConfig& m = c.env.model;
    V3 acc{};
    const double AMS = p.cr * p.area / p.mass;
    bool haveOri = false;
    Orientation ori;
    auto orientation = [&]() -> const Orientation& {
        if (!haveOri) { ori = contextOrientation(c, et); haveOri = true; }
        return ori;
    };
    for (const auto& b : m.bodies) {
        const V3 bodyPos = bodyPosition(c, b.naif, et);
        if (b.pointGravity) {
            if (b.naif == phx::NAIF_EARTH) {
                if (only < 0 || only == 0) acc = acc + twoBody(pos, b.gm);
            } else if (only < 0 || only == 2) {
                acc = acc + thirdBodyNB(pos, bodyPos, b.gm);
            }
        }
        if (b.naif != phx::NAIF_EARTH && (b.shape || b.oblateness)) {
            // SphericalHarmonics.cu:90-144, J2.h:75-90: the body-relative
            // position in the body's axes, nothing beyond its sphere of influence
            if (only < 0 || only == 2) {
                size_t k = 0;
                while (c.env.model.bodies[size_t(c.env.rot[k].body)].naif != b.naif) ++k;
                const V3 d = pos - bodyPos;
                if (!(norm(d) > b.soi)) {
                    const M3 R = rotBodyAxes(c, k, et);
                    if (b.shape) acc = acc + sphericalHarmonics(d, R, c.rotSh[k]);
                    else acc = acc + j2Gravity(d, R, b.gm, b.j2, b.radius, b.soi);
                }
            }
        } else if (b.shape && c.hasSh && (only < 0 || only == 1))
            acc = acc + sphericalHarmonics(pos, orientation().R, c.sh);
        else if (b.oblateness && !b.shape && (only < 0 || only == 1))
            acc = acc + j2Gravity(pos, orientation().R, b.gm, b.j2, b.radius, b.soi);
        if (b.radiation && (only < 0 || only == 3)) {
            const V3 d = pos - bodyPos;
            if (!c.anyOcculter) {
                acc = acc + srp(d, AMS, m.au);
            } else {
                double occ = 1.0;
                for (const auto& o : m.bodies) {
                    if (!o.occulting || o.naif == phx::NAIF_SUN) continue;
                    const V3 occPos = bodyPosition(c, o.naif, et);
                    occ *= shadowFactor(d, pos - occPos, b.radius, o.radius);
                }
                const V3 s = srp(d, AMS, m.au);
                acc = acc + V3{occ * s.x, occ * s.y, occ * s.z};
            }
        }
        if (b.atmosphere && b.nrlmsise && (only < 0 || only == 4)) {
            const Orientation& o = orientation();
            const V3 fixed = mul(o.R, pos);
            const double rho = nrlmsiseDensity(c.nrlTables, c.weather, b.nrl.hCutoff, b.nrl.densityScale, fixed, et,
                                               b.radius, b.flattening);
            const V3 vrel = vel - cross(o.omega, pos);
            const double bc = p.cd * p.area / p.mass;
            const double factor = -0.5 * rho * bc * norm(vrel);
            acc = acc + V3{factor * vrel.x, factor * vrel.y, factor * vrel.z};
        } else if (b.atmosphere && !b.expSegments.empty() && (only < 0 || only == 4)) {
            // PrivateAccelerations.h:167-202: the pole is the body z axis in ICRF
            const Orientation& o = orientation();
            const V3 pole{o.R.m[6], o.R.m[7], o.R.m[8]};
            const double h = geodeticAltitude(pos, pole, b.radius, b.flattening);
            const phx::ExpSegment* bands = b.expSegments.data();
            const int n = int(b.expSegments.size());
            const int band = expBand(bands, n, h);
            if (h <= bands[band].hCutoff) {
                const V3 vrel = vel - cross(o.omega, pos);
                const double rho = expDensity(bands, n, h);
                const double bc = p.cd * p.area / p.mass;
                const double factor = -0.5 * rho * bc * norm(vrel);
                acc = acc + V3{factor * vrel.x, factor * vrel.y, factor * vrel.z};
            }
        }
    }
    if (m.anyCorrections() && (only < 0 || only == 5)) {
        const Orientation* o = m.tidesActive() ? &orientation() : nullptr;
        acc = acc + earthCorrections(c, pos, vel, et, o);
    }
    return acc;
}

// parm/integrate/rk/host/Step.h:109-124 with the stage recursion of :33-105.
void stepFixed(const orc_ctx& c, double y[6], double dt, double epoch0, const Params& p) {
    const Tableau& T = rkf78();
    double k[13][6];
    for (int s = 0; s < 13; ++s) {
        double ys[6];
        for (int d = 0; d < 6; ++d) ys[d] = y[d];
        if (s >= 1) {
            for (int col = 0; col <= s - 1; ++col) {
                const double a = T.a[s - 1][col];
                if (a != 0) {
                    const double coef = dt * a;
                    for (int d = 0; d < 6; ++d) ys[d] += coef * k[col][d];
                }
            }
        }
        double ts = epoch0;
        if (s >= 1) ts += dt * T.c[s - 1];
        const V3 a = acceleration(c, V3{ys[0], ys[1], ys[2]}, V3{ys[3], ys[4], ys[5]}, ts, p);
        k[s][0] = ys[3]; k[s][1] = ys[4]; k[s][2] = ys[5];
        k[s][3] = a.x; k[s][4] = a.y; k[s][5] = a.z;
    }
    double next[6];
    for (int d = 0; d < 6; ++d) next[d] = y[d];
    for (int s = 0; s < 13; ++s) {
        if (T.b[s] != 0) {
            const double cb = dt * T.b[s];
            for (int d = 0; d < 6; ++d) next[d] += cb * k[s][d];
        }
    }
    for (int d = 0; d < 6; ++d) y[d] = next[d];
}

}  // namespace

int orc_acceleration(orc_ctx* ctx, const double* r, const double* v, double et, double mass, double area_km2,
                     double cr_eff, double cd, double* acc_out) {
    return orc_force_term(ctx, -1, r, v, et, mass, area_km2, cr_eff, cd, acc_out);
}

int orc_force_term(orc_ctx* ctx, int which, const double* r, const double* v, double et, double mass,
                   double area_km2, double cr_eff, double cd, double* acc_out) {
    try {
        const V3 a = acceleration(*ctx, V3{r[0], r[1], r[2]}, V3{v[0], v[1], v[2]}, et,
                                  Params{mass, area_km2, cr_eff, cd}, which);
        acc_out[0] = a.x; acc_out[1] = a.y; acc_out[2] = a.z;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int orc_propagate(orc_ctx* ctx, size_t n, const int64_t* id, const double* epoch, const double* state6,
                  const double* mass, const double* drag_area, const double* srp_area, const double* cd,
                  const double* cr, double duration, int threads, double* out_epoch, double* out_state6,
                  orc_stats* stats) {
    try {
        if (n == 0) throw std::runtime_error("targets must be a nonempty array.");
        if (!(duration > 0) || !std::isfinite(duration)) throw std::runtime_error("duration_seconds must be finite and positive.");
        std::set<int64_t> ids;
        std::vector<Params> prm(n);
        for (size_t i = 0; i < n; ++i) {
            if (!ids.insert(id[i]).second) throw std::runtime_error("Target IDs must be unique.");
            for (int d = 0; d < 6; ++d)
                if (!std::isfinite(state6[d * n + i])) throw std::runtime_error("State contains a nonfinite value.");
            if (!std::isfinite(epoch[i]) || !std::isfinite(mass[i]) || !std::isfinite(drag_area[i]) ||
                !std::isfinite(srp_area[i]) || !std::isfinite(cd[i]) || !std::isfinite(cr[i]) || mass[i] <= 0 ||
                drag_area[i] <= 0 || srp_area[i] < 0 || cd[i] < 0 || cr[i] < 0)
                throw std::runtime_error("Invalid target epoch or physical parameter.");
            // driver.cpp:117-120
            prm[i] = Params{mass[i], drag_area[i] * 1e-6, cr[i] * srp_area[i] / drag_area[i], cd[i]};
        }
        const double step = ctx->run.stepSeconds;
        // FixedStep.h:7-29
        std::vector<double> cur(n), end(n), nextDt(n);
        std::vector<char> terminated(n, 0), eos(n, 0);
        long long count = 0;
        for (size_t i = 0; i < n; ++i) {
            const double start = mjdToSpice(epoch[i]);
            cur[i] = start;
            end[i] = mjdToSpice(epoch[i] + duration / 86400.0);
            const double span = end[i] - cur[i];
            if (!(span > 0) || !std::isfinite(span))
                throw std::runtime_error("Duration/step cannot be represented by the integration time/index types.");
            nextDt[i] = std::min(step, span);
            count = std::max(count, (long long)std::ceil(span / step) + 1);
        }
        if (!ctx->env.brot.empty()) {
            const phx::BrotIers& u = ctx->env.brot.iers[0];
            for (size_t i = 0; i < n; ++i)
                if (cur[i] < u.start || end[i] > u.final)
                    throw std::runtime_error("Epoch not in the available orientation range");
        }
        // every evaluated epoch must lie in the ephemeris intervals (Interpolators.h:65-76)
        for (const auto& kv : ctx->ephChains) {
            double s0, s1;
            phx::brieCoverage(ctx->env.brie, kv.second, s0, s1);
            for (size_t i = 0; i < n; ++i)
                if (cur[i] < s0 || end[i] > s1)
                    throw std::runtime_error("Epoch not in the available ephemeris range of body " + std::to_string(kv.first));
        }
        std::vector<double> y(6 * n);
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 6; ++d) y[6 * i + d] = state6[d * n + i];
        const int nth = threads > 0 ? threads : ctx->run.maxThreads;
        double objectSteps = 0;
        const auto t0 = std::chrono::steady_clock::now();
        const long long launches = std::min(count, ctx->run.maxNumSteps);
        for (long long s = 0; s < launches; ++s) {
            long long done = 0;
#pragma omp parallel for num_threads(nth) schedule(static) reduction(+ : done)
            for (long long j = 0; j < (long long)n; ++j) {
                if (terminated[size_t(j)]) continue;
                stepFixed(*ctx, &y[6 * size_t(j)], nextDt[size_t(j)], cur[size_t(j)], prm[size_t(j)]);
                ++done;
                // FixedStep.h:30-37
                cur[size_t(j)] += nextDt[size_t(j)];
                const double remaining = end[size_t(j)] - cur[size_t(j)];
                const bool finished = remaining <= 0;
                terminated[size_t(j)] = finished;
                eos[size_t(j)] = finished;
                nextDt[size_t(j)] = finished ? 0 : std::min(step, remaining);
            }
            objectSteps += double(done);
        }
        // decltype(1.0): the builtin double, also in the FLOP-counting build
        const double secs = std::chrono::duration<decltype(1.0)>(std::chrono::steady_clock::now() - t0).count();
        for (size_t i = 0; i < n; ++i) {
            if (!eos[i]) throw std::runtime_error("A target stopped without reaching the requested end of simulation.");
            out_epoch[i] = spiceToMjd(cur[i]);
            for (int d = 0; d < 6; ++d) out_state6[d * n + i] = y[6 * i + d];
        }
        if (stats) {
            stats->object_steps = objectSteps;
            stats->rhs_evaluations = 13.0 * objectSteps;
            stats->seconds = secs;
            stats->threads = nth;
        }
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int orc_nrlmsise(const char* data_dir, int doy, double sec, double alt, double glat, double glon, double lst,
                 double f107a, double f107, double ap, const double* apa, int use_ap_array, double* d9, double* t2) {
    try {
        static NrlmsiseTables tables;
        static std::string loaded;
        const std::string path = std::string(data_dir) + "/nrlmsise00_coefficients.txt";
        if (loaded != path) { loadNrlmsiseTables(path, tables); loaded = path; }
        nrlmsiseRaw(tables, doy, sec, alt, glat, glon, lst, f107a, f107, ap, apa, use_ap_array, d9, t2);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int orc_body_position(int naif, double et, double* pos3) {
    const V3 p = bodyGeocentric(naif, et);
    pos3[0] = p.x; pos3[1] = p.y; pos3[2] = p.z;
    return std::isfinite(p.x) ? 0 : -1;
}

int orc_earth_orientation(double et, double* R9, double* omega3) {
    const Orientation o = earthOrientation(nullptr, et);
    for (int i = 0; i < 9; ++i) R9[i] = o.R.m[i];
    omega3[0] = o.omega.x; omega3[1] = o.omega.y; omega3[2] = o.omega.z;
    return 0;
}

double orc_geodetic_altitude(const double* r3, const double* pole3, double Req, double f) {
    return geodeticAltitude(V3{r3[0], r3[1], r3[2]}, V3{pole3[0], pole3[1], pole3[2]}, Req, f);
}

double orc_exp_density(const orc_ctx* ctx, double h) {
    for (const auto& b : ctx->env.model.bodies)
        if (b.naif == phx::NAIF_EARTH && b.atmosphere && !b.expSegments.empty())
            return expDensity(b.expSegments.data(), int(b.expSegments.size()), h);
    return std::nan("");
}

int orc_ctx_body_position(const orc_ctx* ctx, int naif, double et, double* pos3) {
    try {
        const V3 p = bodyPosition(*ctx, naif, et);
        pos3[0] = p.x; pos3[1] = p.y; pos3[2] = p.z;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int orc_ctx_earth_orientation(const orc_ctx* ctx, double et, double* R9, double* omega3) {
    const Orientation o = contextOrientation(*ctx, et);
    for (int i = 0; i < 9; ++i) R9[i] = o.R.m[i];
    omega3[0] = o.omega.x; omega3[1] = o.omega.y; omega3[2] = o.omega.z;
    return 0;
}

int orc_body_axes(const orc_ctx* ctx, int naif, double et, double* R9) {
    for (size_t k = 0; k < ctx->env.rot.size(); ++k) {
        if (ctx->env.model.bodies[size_t(ctx->env.rot[k].body)].naif != naif) continue;
        const M3 R = rotBodyAxes(*ctx, k, et);
        for (int i = 0; i < 9; ++i) R9[i] = R.m[i];
        return ctx->env.rot[k].target;
    }
    g_error = "no field of body " + std::to_string(naif) + " in its own axes is configured";
    return -1;
}

// ---------------------------------------------------------------------------
// Test hooks: single functions of the restatement, for the pins in
// tests/test_oracle_pins.py.
// ---------------------------------------------------------------------------

int orc_flop_counts(double* out9) {
#ifdef ORACLE_FLOP_COUNT
    for (int c = 0; c < oracle_flop::kCategories; ++c) {
        out9[c] = f64(oracle_flop::counts.flops[c]);
        out9[oracle_flop::kCategories + c] = f64(oracle_flop::counts.trans[c]);
        out9[2 * oracle_flop::kCategories + c] = f64(oracle_flop::counts.calls[c]);
    }
    return 0;
#else
    (void)out9;
    return -1;
#endif
}

void orc_flop_reset(void) {
#ifdef ORACLE_FLOP_COUNT
    oracle_flop::counts = oracle_flop::Counts{};
    oracle_flop::category = oracle_flop::kPerObject;
#endif
}

int orc_time_at(const orc_ctx* ctx, double et, double* out6) {
    const TimeData t = timeAt(ctx->eop, et);
    out6[0] = t.mjdUtc; out6[1] = t.ut1Utc; out6[2] = t.ttUtc;
    out6[3] = t.xpArcsec; out6[4] = t.ypArcsec; out6[5] = t.valid ? 1.0 : 0.0;
    return t.valid ? 0 : -1;
}

int orc_nrl_inputs(const orc_ctx* ctx, double et, const double* fixed3, double* out) {
    using namespace nrl;
    for (int k = 0; k < 19; ++k) out[k] = std::nan("");
    const double mjdUtc = etToMjdUtc(et);
    out[0] = mjdUtc;
    int year = 0, month = 0, day = 0, hour = 0, minute = 0;
    double second = 0;
    calendarDate(mjdUtc, year, month, day, hour, minute, second);
    out[1] = dayOfYear(year, month, day);
    out[2] = hour * 3600.0 + minute * 60.0 + second;
    out[3] = hpopGastHours(mjdUtc);
    double radius = 6378.137, flat = 0.0033528106647474805;
    for (const auto& b : ctx->env.model.bodies)
        if (b.naif == phx::NAIF_EARTH && b.atmosphere) radius = b.radius, flat = b.flattening;
    double lat = 0, lon = 0, alt = 0;
    if (geodeticFixed(fixed3[0], fixed3[1], fixed3[2], radius, flat, lat, lon, alt)) {
        out[4] = wrapHours(out[3] + lon * kRadToDeg / 15.0);
        out[5] = alt;
        out[6] = lat * kRadToDeg;
        out[7] = lon * kRadToDeg;
    }
    double f107 = 0, f107a = 0, avgAp = 0, ap[7] = {};
    if (!fillWeather(ctx->weather, mjdUtc, f107, f107a, avgAp, ap)) return -1;
    out[8] = f107; out[9] = f107a; out[10] = avgAp;
    for (int k = 0; k < 7; ++k) out[11 + k] = ap[k];
    out[18] = 1.0;
    return 0;
}

namespace {
TideInput tideInput(const double* moon3, const double* sun3, const double* E9, const double* time5) {
    TideInput in{};
    for (int i = 0; i < 3; ++i) { in.moon[i] = moon3[i]; in.sun[i] = sun3[i]; }
    for (int i = 0; i < 9; ++i) in.E[i] = E9[i];
    in.mjdUtc = time5[0]; in.ut1Utc = time5[1]; in.ttUtc = time5[2]; in.xpArcsec = time5[3]; in.ypArcsec = time5[4];
    return in;
}
void unpackCS(const CS7& cs, double* C49, double* S49) {
    for (int n = 0; n < 7; ++n)
        for (int m = 0; m < 7; ++m) { C49[7 * n + m] = cs.C[n][m]; S49[7 * n + m] = cs.S[n][m]; }
}
}  // namespace

int orc_solid_tide(const orc_ctx* ctx, const double* moon3, const double* sun3, const double* E9, const double* time5,
                   int frequency_terms, int zero_tide, double* C49, double* S49) {
    const TideInput in = tideInput(moon3, sun3, E9, time5);
    TideConstants k = ctx->tideK;
    k.zeroTide = zero_tide != 0;
    static const TideTables none;
    CS7 cs;
    solidCoefficients(in, k, frequency_terms ? ctx->tides : none, cs);
    unpackCS(cs, C49, S49);
    return 0;
}

int orc_ocean_waves(const orc_ctx* ctx, const double* args5, double gmst, double* C49, double* S49) {
    CS7 cs;
    clearCS(cs);
    oceanWaves(args5, gmst, ctx->tides, cs);
    unpackCS(cs, C49, S49);
    return 0;
}

int orc_tide_arguments(const double* time5, double* args5, double* gmst, double* beta6) {
    const double zero[9] = {};
    const TideInput in = tideInput(zero, zero, zero, time5);
    tideArguments(in, args5, *gmst);
    doodsonArguments(args5, *gmst, beta6);
    return 0;
}

int orc_third_body(const double* r3, const double* b3, double gm, double* out3) {
    const V3 a = thirdBodyNB(V3{r3[0], r3[1], r3[2]}, V3{b3[0], b3[1], b3[2]}, gm);
    out3[0] = a.x; out3[1] = a.y; out3[2] = a.z;
    return 0;
}

double orc_shadow(const double* sun3, const double* body3, double rSun, double rBody) {
    return shadowFactor(V3{sun3[0], sun3[1], sun3[2]}, V3{body3[0], body3[1], body3[2]}, rSun, rBody);
}

int orc_ocean_tide(const orc_ctx* ctx, const double* time5, double* C49, double* S49) {
    const double zero[9] = {};
    CS7 cs;
    oceanCoefficients(tideInput(zero, zero, zero, time5), ctx->tides, cs);
    unpackCS(cs, C49, S49);
    return 0;
}

int orc_tide_acceleration(const orc_ctx* ctx, const double* r3, const double* E9, const double* C49, const double* S49,
                          int max_degree, double* out3) {
    if (max_degree < 2 || max_degree > 6) { g_error = "max_degree must be 2..6"; return -1; }
    const double zero[9] = {};
    TideInput in = tideInput(zero, zero, E9, zero);
    for (int i = 0; i < 3; ++i) in.r[i] = r3[i];
    CS7 cs;
    for (int n = 0; n < 7; ++n)
        for (int m = 0; m < 7; ++m) { cs.C[n][m] = C49[7 * n + m]; cs.S[n][m] = S49[7 * n + m]; }
    tideAcceleration(in, ctx->tideK, cs, out3, max_degree);
    return 0;
}

int orc_sh_body_fixed(const orc_ctx* ctx, const double* p3, double* out3) {
    if (!ctx->hasSh) { g_error = "no spherical-harmonic field configured"; return -1; }
    const V3 a = shBodyFixed(V3{p3[0], p3[1], p3[2]}, ctx->sh);
    out3[0] = a.x; out3[1] = a.y; out3[2] = a.z;
    return 0;
}
