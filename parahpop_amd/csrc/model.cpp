// CONFIG.json model parsing and environment-table loaders.
// Reference behaviour followed:
//   key checks / path resolution      src/common/driver.cpp:29-55,72-86
//   body feature vocabulary            interface/config/model/environment/bodies/BodyConfig.h:190-262,430-470
//   PDS-SHA gravity file               interface/config/model/sphericalharmonics/Coefficients.h:188-362,480-535
//   CelesTrak SW-All                   model/environment/atmosphere/Nrlmsise00Atmosphere.h:146-236
//   HPOP EOP-All                       model/environment/earthcorrections/EarthCorrections.h:141-177
//   constants.json (SOIOld -> soi)     third_party/brie/brie/constants/ConstantsJson.h:51-74
//   gravity.oblateness, inline j2      interface/config/model/environment/bodies/BodyConfig.h:440-515,
//                                      interface/config/model/environment/Bodies.h:700-750
//   atmosphere.exponential blocks      interface/config/model/environment/Atmosphere.h:130-160,326-350,430-510
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>

namespace phx {

std::string lower(std::string s) {
    for (auto& c : s) c = char(std::tolower((unsigned char)c));
    return s;
}

namespace {

bool isRegularFile(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string resolvePath(const std::string& path, const std::string& base) {
    std::string p = path;
    if (p.empty()) throw std::runtime_error("Empty data path in configuration.");
    if (p[0] != '/') p = base + "/" + p;
    // lexical normalisation of "a/./b" and "a/x/../b"
    std::vector<std::string> parts;
    std::stringstream ss(p);
    std::string item;
    while (std::getline(ss, item, '/')) {
        if (item.empty() || item == ".") continue;
        if (item == "..") { if (!parts.empty()) parts.pop_back(); continue; }
        parts.push_back(item);
    }
    std::string out;
    for (auto& s : parts) out += "/" + s;
    if (out.empty()) out = "/";
    if (!isRegularFile(out)) throw std::runtime_error("Required data file not found: " + out);
    return out;
}

void checkKeys(const json::Value& v, std::initializer_list<const char*> allowed, const char* ctx) {
    if (!v.isObject()) throw std::runtime_error(std::string(ctx) + " must be an object.");
    for (const auto& kv : *v.obj) {
        bool ok = false;
        for (auto* k : allowed) ok |= kv.first == k;
        if (!ok) throw std::runtime_error(std::string("Unsupported ") + ctx + " option: " + kv.first);
    }
}

struct BodyName { const char* key; int naif; };
const BodyName kBodyNames[] = {
    {"sun", NAIF_SUN}, {"moon", NAIF_MOON}, {"earth", NAIF_EARTH},
    {"mercurybarycenter", NAIF_MERCURY_BARY}, {"venusbarycenter", NAIF_VENUS_BARY},
    {"marsbarycenter", NAIF_MARS_BARY}, {"jupiterbarycenter", NAIF_JUPITER_BARY},
    {"saturnbarycenter", NAIF_SATURN_BARY}, {"uranusbarycenter", NAIF_URANUS_BARY},
    {"neptunebarycenter", NAIF_NEPTUNE_BARY}, {"plutobarycenter", NAIF_PLUTO_BARY},
};

int naifFromName(const std::string& name) {
    std::string k;
    for (char c : lower(name)) if (c != ' ' && c != '_' && c != '-') k += c;
    for (const auto& b : kBodyNames) if (k == b.key) return b.naif;
    throw std::runtime_error("Unsupported body: " + name +
        " (the analytic ephemeris provides the Sun, Moon, Earth and the planetary system barycenters)");
}

bool activationValue(const json::Value& v) {
    if (v.isBool()) return v.b;
    if (v.isNull()) return false;
    if (v.isObject() || v.isArray()) return true;
    if (v.isString()) {
        if (lower(v.str) == "default") return true;
        throw std::runtime_error("Invalid feature value: " + v.str);
    }
    throw std::runtime_error("Unsupported JSON type for a body feature.");
}

void parseShape(BodySpec& b, const json::Value& s, const std::string& base) {
    if (!s.isObject()) throw std::runtime_error("gravity.shape must be an object with a coefficient file.");
    for (const auto& kv : *s.obj) {
        const std::string k = lower(kv.first);
        if (k == "body") {
            if (naifFromName(kv.second.asString()) != b.naif)
                throw std::runtime_error("gravity.shape.body does not match the enclosing body.");
        } else if (k == "file") {
            b.shFile = resolvePath(kv.second.asString(), base);
        } else if (k == "delimiter") {
            const std::string& d = kv.second.asString();
            if (d.size() != 1) throw std::runtime_error("gravity.shape.delimiter must be one character.");
            b.shDelimiter = d[0];
        } else if (k == "headerlines") {
            b.shHeaderLines = int(kv.second.asInt());
        } else if (k == "degree") {
            b.shDegree = int(kv.second.asInt());
        } else if (k == "order") {
            b.shOrder = int(kv.second.asInt());
        } else {
            throw std::runtime_error("Unsupported gravity.shape option: " + kv.first);
        }
    }
    if (b.shFile.empty()) throw std::runtime_error("gravity.shape requires a coefficient file.");
}

void parseGravity(BodySpec& b, const json::Value& g, const std::string& base) {
    if (g.isBool()) {
        b.pointGravity = g.b;
        if (!g.b) b.shape = false;
    } else if (g.isNull()) {
        b.pointGravity = b.shape = false;
    } else if (g.isString()) {
        if (lower(g.str) != "default") throw std::runtime_error("Invalid gravity value string: " + g.str);
        b.pointGravity = true;
    } else if (g.isObject()) {
        for (const auto& kv : *g.obj) {
            const std::string k = lower(kv.first);
            if (k == "shape") {
                b.shape = activationValue(kv.second);
                if (b.shape) parseShape(b, kv.second, base);
            } else if (k == "oblateness") {
                b.oblateness = activationValue(kv.second);
            } else {
                throw std::runtime_error("Unsupported gravity feature: " + kv.first);
            }
        }
    } else if (g.isArray()) {
        for (const auto& item : *g.arr) {
            const std::string k = lower(item.asString());
            if (k == "oblateness") { b.oblateness = true; continue; }
            if (k == "shape") throw std::runtime_error("gravity shape requires a coefficient-file object.");
            throw std::runtime_error("Unsupported gravity feature: " + item.str);
        }
    } else {
        throw std::runtime_error("Unsupported JSON type for gravity features.");
    }
}

// One {rho0, h0, scaleHeight, hCutoff} block: all four required, no other key
// (Atmosphere.h:461-505).
ExpSegment parseExpSegment(const json::Value& v) {
    if (!v.isObject()) throw std::runtime_error("Each exponential atmosphere block must be an object.");
    ExpSegment s;
    bool rho0 = false, h0 = false, scale = false, cutoff = false;
    for (const auto& kv : *v.obj) {
        const std::string k = lower(kv.first);
        if (k == "rho0") { s.rho0 = kv.second.asDouble(); rho0 = true; }
        else if (k == "h0") { s.h0 = kv.second.asDouble(); h0 = true; }
        else if (k == "scaleheight") { s.scaleHeight = kv.second.asDouble(); scale = true; }
        else if (k == "hcutoff") { s.hCutoff = kv.second.asDouble(); cutoff = true; }
        else throw std::runtime_error("Invalid atmosphere field: '" + kv.first +
                                      "'. Valid fields are: [h0, hcutoff, rho0, scaleheight] (case insensitive)");
    }
    if (!rho0) throw std::runtime_error("Atmosphere rho0 not specified.");
    if (!h0) throw std::runtime_error("Atmosphere h0 not specified.");
    if (!scale) throw std::runtime_error("Atmosphere scaleHeight not specified.");
    if (!cutoff) throw std::runtime_error("Atmosphere hCutoff not specified.");
    if (!(s.rho0 >= 0.0)) throw std::runtime_error("Atmosphere rho0 must be non-negative.");
    if (!(s.scaleHeight > 0.0)) throw std::runtime_error("Atmosphere scaleHeight must be positive.");
    return s;
}

// A single block or a list of blocks, sorted ascending by h0; a band may not
// start below the previous band's cutoff (a gap reads as vacuum) (Atmosphere.h:326-350,430-459).
void parseExponential(BodySpec& b, const json::Value& v) {
    std::vector<ExpSegment> segs;
    if (v.isArray()) {
        if (v.size() == 0) throw std::runtime_error("Piecewise atmosphere list must contain at least one block.");
        for (const auto& e : *v.arr) segs.push_back(parseExpSegment(e));
    } else if (v.isObject()) {
        segs.push_back(parseExpSegment(v));
    } else {
        throw std::runtime_error("Exponential atmosphere must be an object (single block) or a list of objects (piecewise).");
    }
    std::stable_sort(segs.begin(), segs.end(), [](const ExpSegment& x, const ExpSegment& y) { return x.h0 < y.h0; });
    for (size_t i = 1; i < segs.size(); ++i)
        if (!(segs[i].h0 >= segs[i - 1].hCutoff))
            throw std::runtime_error("Overlapping atmosphere segments: a band with h0=" + std::to_string(segs[i].h0) +
                                     " km starts below the previous band's hCutoff=" + std::to_string(segs[i - 1].hCutoff) + " km.");
    b.expSegments = std::move(segs);
}

void parseAtmosphere(BodySpec& b, const json::Value& a, const std::string& base) {
    b.atmosphere = activationValue(a);
    if (!a.isObject()) {
        if (b.atmosphere) throw std::runtime_error("atmosphere requires a model block (exponential or nrlmsise00).");
        return;
    }
    if (a.size() != 1)
        throw std::runtime_error("Atmosphere block must be a single-key dispatch dict naming the model (exponential or nrlmsise00).");
    for (const auto& kv : *a.obj) {
        const std::string k = lower(kv.first);
        if (k == "exponential") {
            parseExponential(b, kv.second);
            continue;
        }
        if (k != "nrlmsise00" && k != "nrlmsise-00")
            throw std::runtime_error("Unknown atmosphere model: '" + kv.first + "'.  Supported: exponential, nrlmsise00.");
        b.nrlmsise = true;
        const json::Value& n = kv.second;
        if (!n.isObject()) throw std::runtime_error("nrlmsise00 must be an object.");
        for (const auto& p : *n.obj) {
            const std::string q = lower(p.first);
            if (q == "spaceweatherfile") b.nrl.spaceWeatherFile = resolvePath(p.second.asString(), base);
            else if (q == "coveragestartmjdutc") b.nrl.coverageStartMjdUtc = p.second.asDouble();
            else if (q == "coverageendmjdutc") b.nrl.coverageEndMjdUtc = p.second.asDouble();
            else if (q == "hcutoff") b.nrl.hCutoff = p.second.asDouble();
            else if (q == "densityscale") b.nrl.densityScale = p.second.asDouble();
            else throw std::runtime_error("Unsupported nrlmsise00 option: " + p.first);
        }
        if (b.nrl.spaceWeatherFile.empty()) throw std::runtime_error("nrlmsise00 requires spaceWeatherFile.");
        if (!(b.nrl.hCutoff > 0) || !(b.nrl.densityScale >= 0))
            throw std::runtime_error("nrlmsise00 hCutoff must be positive and densityScale non-negative.");
    }
}

struct ConstEntry { int naif; double gm, r, soi, j2; bool hasGm; };

void loadConstants(const std::string& path, ModelConfig& m, std::vector<ConstEntry>& out) {
    json::Value root = json::readFile(path);
    if (!root.isArray()) throw std::runtime_error("constants.json must be an array.");
    for (const auto& e : *root.arr) {
        if (!e.isObject()) continue;
        if (const auto* v = e.find("AU")) m.au = v->asDouble();
        if (const auto* v = e.find("CLIGHT")) m.clight = v->asDouble();
        if (const auto* v = e.find("EMRAT")) m.emrat = v->asDouble();
        const auto* id = e.find("NAIFID");
        if (!id) continue;
        ConstEntry c{int(id->asDouble()), 0.0, 0.0, 1e20, 0.0, false};
        if (const auto* v = e.find("GM")) { c.gm = v->asDouble(); c.hasGm = true; }
        if (const auto* v = e.find("R")) c.r = v->asDouble();
        if (const auto* v = e.find("SOIOld")) c.soi = v->asDouble();
        if (const auto* v = e.find("J2")) c.j2 = v->asDouble();
        out.push_back(c);
    }
}

}  // namespace

void parseConfig(const json::Value& root, const std::string& configDir,
                 RunConfig& run, ModelConfig& model, bool requireRun) {
    checkKeys(root, {"run", "execution", "model"}, "configuration");
    if (requireRun || root.has("run")) {
        checkKeys(root.at("run"), {"duration_seconds"}, "run");
        run.durationSeconds = root.at("run").at("duration_seconds").asDouble();
        if (!(run.durationSeconds > 0) || !std::isfinite(run.durationSeconds))
            throw std::runtime_error("duration_seconds must be finite and positive.");
    }
    const json::Value& ex = root.at("execution");
    checkKeys(ex, {"maxThreads", "integration", "precision"}, "execution");
    checkKeys(ex.at("integration"), {"initialStepSize", "maxNumSteps"}, "integration");
    run.maxNumSteps = ex.at("integration").at("maxNumSteps").asInt();
    run.maxThreads = ex.has("maxThreads") ? int(ex.at("maxThreads").asInt()) : 1;
    if (run.maxNumSteps <= 0 || run.maxThreads <= 0)
        throw std::runtime_error("maxNumSteps and maxThreads must be positive integers.");
    run.stepSeconds = ex.at("integration").at("initialStepSize").asDouble();
    if (!(run.stepSeconds > 0) || !std::isfinite(run.stepSeconds))
        throw std::runtime_error("initialStepSize must be finite and positive.");
    if (const auto* p = ex.find("precision")) {
        checkKeys(*p, {"mode", "fp64Degree", "fp64Order"}, "precision");
        const std::string mode = lower(p->at("mode").asString());
        if (mode == "mixed") run.precision.mixed = true;
        else if (mode != "fp64") throw std::runtime_error("precision.mode must be \"fp64\" or \"mixed\".");
        if (p->has("fp64Degree")) run.precision.fp64Degree = int(p->at("fp64Degree").asInt());
        if (run.precision.fp64Degree < 1) throw std::runtime_error("precision.fp64Degree must be >= 1.");
        if (p->has("fp64Order")) run.precision.fp64Order = int(p->at("fp64Order").asInt());
        if (run.precision.fp64Order < 0) throw std::runtime_error("precision.fp64Order must be >= 0.");
    }

    const json::Value& env = root.at("model").at("environment");
    checkKeys(root.at("model"), {"environment"}, "model");
    if (!env.isObject()) throw std::runtime_error("model.environment must be an object.");
    std::vector<ConstEntry> consts;
    for (const auto& kv : *env.obj) {
        const std::string k = lower(kv.first);
        if (k == "bodies") continue;
        if (k == "constants") {
            model.constantsFile = resolvePath(kv.second.asString(), configDir);
        } else if (k == "ephemeris") {
            // "analytic", or one or more .brie files (Environment.h:38; EphUnit.h:508-560)
            std::vector<std::string> v;
            if (kv.second.isString()) v.push_back(kv.second.str);
            else if (kv.second.isArray() && kv.second.size() > 0)
                for (const auto& e : *kv.second.arr) v.push_back(e.asString());
            else throw std::runtime_error("ephemeris must name \"analytic\" or a list of .brie files.");
            if (v.size() == 1 && lower(v[0]) == "analytic") {
                model.ephemeris = "analytic";
            } else {
                model.ephemeris = "brie";
                for (const auto& f : v) model.ephemerisFiles.push_back(resolvePath(f, configDir));
            }
        } else if (k == "orientations") {
            // "analytic", or .brot2 files holding an IERS2000 unit (orientations/EphUnit.h:405-445)
            std::vector<std::string> v;
            if (kv.second.isString()) v.push_back(kv.second.str);
            else if (kv.second.isArray() && kv.second.size() > 0)
                for (const auto& e : *kv.second.arr) v.push_back(e.asString());
            else throw std::runtime_error("orientations must name \"analytic\" or a list of .brot2 files.");
            if (v.size() == 1 && lower(v[0]) == "analytic") {
                model.orientation = "analytic";
            } else {
                model.orientation = "brot2";
                for (const auto& f : v) model.orientationFiles.push_back(resolvePath(f, configDir));
            }
        } else if (k == "prefserssb" || k == "preferssb") {
            model.preferSSB = kv.second.asBool();
        } else if (k == "earthcorrections") {
            const json::Value& ec = kv.second;
            if (!ec.isObject()) throw std::runtime_error("earthCorrections must be an object.");
            for (const auto& p : *ec.obj) {
                const std::string q = lower(p.first);
                if (q == "model") {
                    if (p.second.asString() != "iers2010_fes2004")
                        throw std::runtime_error("earthCorrections.model must be iers2010_fes2004.");
                } else if (q == "solidearthtides") model.solidTides = p.second.asBool();
                else if (q == "oceantides") model.oceanTides = p.second.asBool();
                else if (q == "relativity") model.relativity = p.second.asBool();
                else if (q == "tidesystem") {
                    const std::string t = p.second.asString();
                    if (t != "zero_tide" && t != "tide_free")
                        throw std::runtime_error("tideSystem must be zero_tide or tide_free.");
                    model.zeroTide = t == "zero_tide";
                } else if (q == "eopfile") model.eopFile = resolvePath(p.second.asString(), configDir);
                else throw std::runtime_error("Unsupported earthCorrections option: " + p.first);
            }
            if (model.tidesActive() && model.eopFile.empty())
                throw std::runtime_error("Earth tides require earthCorrections.eopFile.");
        } else {
            throw std::runtime_error("Unsupported model.environment option: " + kv.first);
        }
    }
    if (!model.constantsFile.empty()) loadConstants(model.constantsFile, model, consts);

    const json::Value& bodies = env.at("bodies");
    if (!bodies.isObject() || bodies.size() == 0)
        throw std::runtime_error("model.environment.bodies must be a nonempty object.");
    for (const auto& kv : *bodies.obj) {
        BodySpec b;
        b.name = kv.first;
        b.naif = naifFromName(kv.first);
        for (const auto& c : consts)
            if (c.naif == b.naif) { if (c.hasGm) b.gm = c.gm; b.radius = c.r; b.soi = c.soi; b.j2 = c.j2; }
        const json::Value& f = kv.second;
        if (f.isObject()) {
            for (const auto& p : *f.obj) {
                const std::string q = lower(p.first);
                if (q == "gravity") parseGravity(b, p.second, configDir);
                else if (q == "radiation") b.radiation = activationValue(p.second);
                else if (q == "atmosphere") parseAtmosphere(b, p.second, configDir);
                else if (q == "occulting") b.occulting = activationValue(p.second);
                else if (q == "flattening") {
                    b.flattening = p.second.asDouble();
                    if (!(b.flattening >= 0.0 && b.flattening < 1.0))
                        throw std::runtime_error("flattening must lie in [0, 1) for body " + b.name);
                }
                else if (q == "gm") b.gm = p.second.asDouble();
                else if (q == "radius") b.radius = p.second.asDouble();
                else if (q == "soi") b.soi = p.second.asDouble();
                else if (q == "j2") b.j2 = p.second.asDouble();
                else if (q == "orientation") b.orientation = p.second.asString();  // Earth uses the IERS2000 unit
                else if (q == "amean") (void)p.second.asDouble();  // accepted; not used by these forces
                else throw std::runtime_error("Unsupported body option: " + p.first);
            }
        } else if (f.isArray()) {
            for (const auto& item : *f.arr) {
                const std::string q = lower(item.asString());
                if (q == "gravity") b.pointGravity = true;
                else if (q == "radiation") b.radiation = true;
                else if (q == "occulting") b.occulting = true;
                else throw std::runtime_error("Unsupported body feature: " + item.str);
            }
        } else if (f.isBool() || f.isNull()) {
            if (!activationValue(f)) b.pointGravity = false;
        } else if (!(f.isString() && lower(f.str) == "default")) {
            throw std::runtime_error("Unsupported JSON type for body features.");
        }
        if (model.find(b.naif) >= 0) throw std::runtime_error("Body listed twice: " + kv.first);
        model.bodies.push_back(b);
    }

    // Validation that mirrors the reference model build (physics/reduced/Dimensional.h:49-144).
    if (model.find(NAIF_EARTH) < 0) throw std::runtime_error("Earth must be configured (it is the centre of integration).");
    for (const auto& b : model.bodies) {
        if ((b.pointGravity || b.shape) && !(b.gm > 0))
            throw std::runtime_error("No gravitational parameter for body " + b.name);
        if (b.radiation && b.naif != NAIF_SUN) throw std::runtime_error("Radiation is supported for the Sun only.");
        if (b.atmosphere && b.naif != NAIF_EARTH) throw std::runtime_error("Atmospheres are supported for Earth only.");
        if ((b.shape || b.oblateness) && b.naif != NAIF_EARTH && !(b.radius > 0))
            throw std::runtime_error("Body field needs the body radius: " + b.name);
        if (b.atmosphere && !b.nrlmsise && !(b.radius > 0))
            throw std::runtime_error("Exponential atmosphere needs the body radius: " + b.name);
        if (b.occulting && !(b.radius > 0)) throw std::runtime_error("Occulting body needs a radius: " + b.name);
        if (b.radiation && !(b.radius > 0)) throw std::runtime_error("Sun radius missing for shadow geometry.");
    }
    int rotating = 0;
    for (const auto& b : model.bodies) rotating += (b.shape || b.oblateness) && b.naif != NAIF_EARTH ? 1 : 0;
    if (rotating > kMaxRotBodies)
        throw std::runtime_error("At most " + std::to_string(kMaxRotBodies) +
                                 " bodies other than the Earth may carry a gravity field.");
    if (rotating > 0 && model.orientationFiles.empty())
        throw std::runtime_error("A gravity field of a body other than the Earth needs its rotation data: list the "
                                 "defining .brot2 files under model.environment.orientations.");
}

ShField loadShField(const std::string& path, char delimiter, int headerLines, int degree, int order) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open file: " + path);
    std::vector<std::vector<double>> rows;
    std::string line;
    while (std::getline(in, line)) {
        if (headerLines > 0) { --headerLines; continue; }
        line.erase(std::find(line.begin(), line.end(), '#'), line.end());
        line.erase(std::remove_if(line.begin(), line.end(), [](char c) { return c == '\n' || c == '\r'; }), line.end());
        if (line.empty()) continue;
        rows.emplace_back();
        std::string tok;
        std::stringstream ss(line);
        while (std::getline(ss, tok, delimiter)) {
            size_t a = tok.find_first_not_of(" \t"), b = tok.find_last_not_of(" \t");
            if (a == std::string::npos) continue;
            tok = tok.substr(a, b - a + 1);
            // Fortran-style exponents (1.0D-05) are accepted as in the reference's strtod-based parse
            for (auto& c : tok) if (c == 'D' || c == 'd') c = 'e';
            char* end = nullptr;
            double v = std::strtod(tok.c_str(), &end);
            if (end == tok.c_str()) throw std::runtime_error("Invalid number '" + tok + "' in file: " + path);
            rows.back().push_back(v);
        }
    }
    if (rows.empty()) throw std::runtime_error("No coefficients found in file: " + path);
    const auto& h = rows.front();
    if (h.size() != 8)
        throw std::runtime_error("Invalid PDS-SHA header in file: " + path + " - Expected 8 values, found " + std::to_string(h.size()));
    const int degH = int(h[3]), ordH = int(h[4]), normH = int(h[5]);
    if (degree != -1 && degree > degH) throw std::runtime_error("Requested maximum degree exceeds available maximum degree in file: " + path);
    if (order != -1 && order > ordH) throw std::runtime_error("Requested maximum order exceeds available maximum order in file: " + path);
    if (normH != 0 && normH != 1) throw std::runtime_error("Unsupported normalization flag in file: " + path);
    if (h[6] != 0.0 || h[7] != 0.0) throw std::runtime_error("Non-zero latitude/longitude offsets are not supported.");
    ShField f;
    f.degree = degree == -1 ? degH : degree;
    f.order = order == -1 ? ordH : order;
    if (f.order > f.degree) throw std::runtime_error("Spherical-harmonic order exceeds degree.");
    if (f.degree < 1) throw std::runtime_error("Maximum degree must be > 0");
    f.normalized = normH == 1;
    f.radius = h[0];
    f.gm = h[1];
    const size_t n = size_t(f.degree + 1) * (f.degree + 2) / 2;
    f.C.assign(n, 0.0);
    f.S.assign(n, 0.0);
    for (size_t i = 1; i < rows.size(); ++i) {
        const auto& r = rows[i];
        if (r.size() != 4 && r.size() != 6)
            throw std::runtime_error("Invalid coefficient row in file: " + path);
        const int dn = int(r[0]), dm = int(r[1]);
        if (dn <= 0) throw std::runtime_error("Invalid degree 0 coefficient found in file: " + path);
        if (dm > dn) throw std::runtime_error("Invalid order for degree in file: " + path);
        if (dm == 0 && r[3] != 0.0) throw std::runtime_error("Invalid non-zero Sn0 coefficient found in file: " + path);
        if (dn > f.degree || dm > f.order) continue;
        f.C[f.index(dn, dm)] = r[2];
        if (dm > 0) f.S[f.index(dn, dm)] = r[3];
    }
    shFactors(f);
    return f;
}

void shFactors(ShField& f) {
    const int L = f.degree;
    f.factors.assign(size_t(L + 1) * (L + 2), 0.0);
    for (int m = 0; m <= f.order; ++m) {
        for (int n = m; n <= L; ++n) {
            const size_t idx = 2 * f.index(n, m);
            if (n == m) {
                double cm;
                if (f.normalized) cm = m == 0 ? 1.0 : (m == 1 ? std::sqrt(3.0) : std::sqrt((2.0 * m + 1.0) / (2.0 * m)));
                else cm = m <= 1 ? 1.0 : 2.0 * m - 1.0;
                f.factors[idx] = cm;
                f.factors[idx + 1] = 0.0;
            } else {
                const double nn = n, mm = m;
                double a, b;
                if (f.normalized) {
                    a = std::sqrt(((2.0 * nn - 1.0) * (2.0 * nn + 1.0)) / ((nn - mm) * (nn + mm)));
                    b = std::sqrt(((2.0 * nn + 1.0) * (nn + mm - 1.0) * (nn - mm - 1.0)) /
                                  ((2.0 * nn - 3.0) * (nn + mm) * (nn - mm)));
                } else {
                    a = (2.0 * nn - 1.0) / (nn - mm);
                    b = (nn + mm - 1.0) / (nn - mm);
                }
                f.factors[idx] = a;
                f.factors[idx + 1] = b;
            }
        }
    }
}

double calendarToMjd(int year, int month, int day) {
    const int a = (14 - month) / 12;
    const int y = year + 4800 - a;
    const int m = month + 12 * a - 3;
    const long jdn = day + (153 * m + 2) / 5 + 365L * y + y / 4 - y / 100 + y / 400 - 32045;
    return double(jdn) - 2400001.0;
}

WeatherTable loadSpaceWeather(const std::string& path, double requestedStart, double requestedEnd) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Unable to open NRLMSISE-00 space-weather file: " + path);
    struct Rec { double mjd; int ap[8]; int avgAp; double obsF, obsFA, adjF, adjFA; };
    std::vector<Rec> all;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream s(line);
        int year, month, day, bsrn, nd, kp[8], sumKp, c9, isn, q;
        double cp, adjLast, obsLast;
        Rec r{};
        if (!(s >> year >> month >> day >> bsrn >> nd)) continue;
        bool ok = true;
        for (int& v : kp) ok = ok && bool(s >> v);
        ok = ok && bool(s >> sumKp);
        for (int& v : r.ap) ok = ok && bool(s >> v);
        ok = ok && bool(s >> r.avgAp >> cp >> c9 >> isn >> r.adjF >> q >> r.adjFA >> adjLast >> r.obsF >> r.obsFA >> obsLast);
        if (!ok) continue;
        r.mjd = calendarToMjd(year, month, day);
        all.push_back(r);
    }
    if (all.empty()) throw std::runtime_error("No CelesTrak SW-All records found in: " + path);
    const double keepStart = requestedStart >= 0.0 ? std::floor(requestedStart) - 4.0 : all.front().mjd;
    const double keepEnd = requestedEnd >= 0.0 ? std::floor(requestedEnd) + 1.0 : all.back().mjd;
    std::vector<Rec> sel;
    for (const auto& r : all) if (r.mjd >= keepStart && r.mjd <= keepEnd) sel.push_back(r);
    if (sel.size() < 5)
        throw std::runtime_error("NRLMSISE-00 weather coverage is missing or too short; at least four look-back days plus the propagation day are required.");
    for (size_t i = 1; i < sel.size(); ++i)
        if (sel[i].mjd != sel[i - 1].mjd + 1.0)
            throw std::runtime_error("NRLMSISE-00 space-weather records must be consecutive days.");
    WeatherTable w;
    w.startMjdUtc = sel.front().mjd;
    w.days = int(sel.size());
    w.data.assign(size_t(WeatherTable::NCOMP) * w.days, 0.0);
    for (int i = 0; i < w.days; ++i) {
        w.data[size_t(WeatherTable::OBS_F107) * w.days + i] = sel[i].obsF;
        w.data[size_t(WeatherTable::OBS_F107A) * w.days + i] = sel[i].obsFA;
        w.data[size_t(WeatherTable::ADJ_F107) * w.days + i] = sel[i].adjF;
        w.data[size_t(WeatherTable::ADJ_F107A) * w.days + i] = sel[i].adjFA;
        w.data[size_t(WeatherTable::AVG_AP) * w.days + i] = sel[i].avgAp;
        for (int k = 0; k < 8; ++k) w.data[size_t(WeatherTable::AP0 + k) * w.days + i] = sel[i].ap[k];
    }
    return w;
}

EopTable loadEop(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open Earth-tide EOP file: " + path);
    EopTable t;
    std::string line;
    double prev = 0.0;
    while (std::getline(in, line)) {
        std::istringstream s(line);
        int year, month, day;
        double mjd, xp, yp, dut, lod, dpsi, deps, dx, dy, dat;
        if (!(s >> year >> month >> day >> mjd >> xp >> yp >> dut >> lod >> dpsi >> deps >> dx >> dy >> dat)) continue;
        if (!(year >= 1960 && month >= 1 && month <= 12 && day >= 1 && day <= 31 && std::isfinite(mjd) &&
              std::isfinite(xp) && std::isfinite(yp) && std::isfinite(dut) && std::isfinite(dat) && mjd == std::floor(mjd)))
            throw std::runtime_error("Invalid Earth-tide EOP record in: " + path);
        if (t.days > 0 && mjd != prev + 1.0)
            throw std::runtime_error("Earth-tide EOP records must be sorted, unique, consecutive days: " + path);
        if (t.days == 0) t.firstMjdUtc = mjd;
        prev = mjd;
        t.xp.push_back(xp); t.yp.push_back(yp); t.dut.push_back(dut); t.dat.push_back(dat);
        ++t.days;
    }
    if (t.days < 2) throw std::runtime_error("Earth tides require at least two HPOP EOP-All records in: " + path);
    return t;
}

namespace {

// Rotation-data target of a body frame name (naming/orientations/Registry.h:
// 580-640 builtins; <BODY>_BODY_FIXED and IAU_<BODY> name a body's IAU model).
// Sets `body` to the body the frame describes.
int frameTarget(const std::string& name, int& body) {
    std::string k;
    for (char c : name) if (c != ' ') k += char(std::toupper((unsigned char)c));
    for (const char* inertial : {"ICRF", "J2000", "EME2000", "ECLIPJ2000", "ECLIPTIC", "ECLIPTICJ2000"})
        if (k == inertial)
            throw std::runtime_error("Body orientation '" + name + "' is an inertial frame; a body's force-model "
                                     "rotation needs a body-fixed or binary frame");
    if (k == "MOON_PA") { body = NAIF_MOON; return 31008; }
    if (k == "ITRF93" || k == "ITRF" || k == "EARTH_FIXED") { body = NAIF_EARTH; return 3000; }
    if (k == "ITRF_IERS2000" || k == "ITRF_IPF" || k == "ITRF_NATIVE") { body = NAIF_EARTH; return 3001; }
    std::string token;
    const std::string suffix = "_BODY_FIXED";
    if (k.size() > suffix.size() && k.compare(k.size() - suffix.size(), suffix.size(), suffix) == 0)
        token = k.substr(0, k.size() - suffix.size());
    else if (k.size() > 4 && k.compare(0, 4, "IAU_") == 0)
        token = k.substr(4);
    if (!token.empty()) {
        try {
            body = naifFromName(token);
            return body;
        } catch (const std::runtime_error&) {
        }
    }
    throw std::runtime_error("Body orientation '" + name + "' is not a known orientation frame; name a builtin "
                             "(MOON_PA, <BODY>_BODY_FIXED, IAU_<BODY>)");
}

// The rotation target of a body's field (naming/orientations/Resolver.h:45-92):
// an explicit frame name, else the unique loaded binary frame of the body,
// else the body's IAU model (its own NAIF id).
void resolveRotBody(const Environment& env, const BodySpec& b, RotBody& r) {
    const BrotTable& t = env.brot;
    if (!b.orientation.empty()) {
        int body = 0;
        r.target = frameTarget(b.orientation, body);
        if (body != b.naif)
            throw std::runtime_error("Body orientation '" + b.orientation + "' describes body " + std::to_string(body) +
                                     ", not body " + std::to_string(b.naif) + "; set it on the body it belongs to");
    } else {
        r.target = b.naif;
        if (b.naif == NAIF_MOON && t.findBody(31008) >= 0) r.target = 31008;
    }
    r.unit = t.findBody(r.target);
    if (r.unit < 0 || (t.bodies[size_t(r.unit)].type != 1 && t.bodies[size_t(r.unit)].type != 3))
        throw std::runtime_error("Body orientation of " + b.name + " (rotation target " + std::to_string(r.target) +
                                 ") needs its rotation data loaded; add the defining .brot2 to the environment's "
                                 "orientations");
    const BrotBody& u = t.bodies[size_t(r.unit)];
    r.bary = -1;
    if (u.type == 1) {
        const int bary = t.findBody(u.p[4]);
        if (bary >= 0) {
            if (t.bodies[size_t(bary)].type != 2)
                throw std::runtime_error("Barycenter " + std::to_string(u.p[4]) + " of body orientation " +
                                         std::to_string(r.target) + " is not a nutation-precession unit");
            const BrotBody& bu = t.bodies[size_t(bary)];
            if (bu.p[1] < std::max(u.p[5], std::max(u.p[6], u.p[7])))
                throw std::runtime_error("Barycenter unit " + std::to_string(u.p[4]) + " holds fewer angles than body "
                                         "orientation " + std::to_string(r.target) + " has terms");
            r.bary = bary;
        }
    }
}

}  // namespace

void loadEnvironment(Environment& env) {
    const int e = env.model.find(NAIF_EARTH);
    const BodySpec& earth = env.model.bodies[size_t(e)];
    if (earth.shape) {
        env.sh = loadShField(earth.shFile, earth.shDelimiter, earth.shHeaderLines, earth.shDegree, earth.shOrder);
        env.hasSh = true;
    }
    if (earth.atmosphere && earth.nrlmsise) {
        env.weather = loadSpaceWeather(earth.nrl.spaceWeatherFile, earth.nrl.coverageStartMjdUtc, earth.nrl.coverageEndMjdUtc);
        env.hasWeather = true;
    }
    if (env.model.tidesActive()) {
        env.eop = loadEop(env.model.eopFile);
        env.hasEop = true;
    }
    if (env.model.orientation == "brot2") {
        for (const auto& f : env.model.orientationFiles) loadBrot2(f, env.brot);
        // the Earth's IERS2000 unit, when one is loaded (the analytic model otherwise)
        if (env.brot.iers.size() > 1)
            throw std::runtime_error("orientations: expected at most one IERS2000 Earth orientation unit, found " +
                                     std::to_string(env.brot.iers.size()));
    }
    env.rot.clear();
    for (size_t i = 0; i < env.model.bodies.size(); ++i) {
        const BodySpec& b = env.model.bodies[i];
        if (b.naif == NAIF_EARTH || !(b.shape || b.oblateness)) continue;
        RotBody r;
        r.body = int(i);
        resolveRotBody(env, b, r);
        r.sh = b.shape;
        if (b.shape) r.field = loadShField(b.shFile, b.shDelimiter, b.shHeaderLines, b.shDegree, b.shOrder);
        else if (!(b.j2 != 0.0)) throw std::runtime_error("No J2 for body " + b.name);
        env.rot.push_back(std::move(r));
    }
    if (env.model.ephemeris == "brie") {
        for (const auto& f : env.model.ephemerisFiles) loadBrie(f, env.brie);
        // every configured body (and the Sun and Moon of the tide model) must
        // chain to the Earth
        for (const auto& b : env.model.bodies) (void)brieChain(env.brie, b.naif, NAIF_EARTH);
        if (env.model.tidesActive()) {
            (void)brieChain(env.brie, NAIF_SUN, NAIF_EARTH);
            (void)brieChain(env.brie, NAIF_MOON, NAIF_EARTH);
        }
    }
}

}  // namespace phx
