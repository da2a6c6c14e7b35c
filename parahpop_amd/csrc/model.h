// Host-side description of one propagation run: the force-model
// configuration read from CONFIG.json ("model.environment") and the
// environment tables it names (gravity field, space weather, EOP, constants).
// Everything here is plain C++ and plain arrays; device upload happens in the
// C-ABI layer.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "brie.h"
#include "json_lite.h"
#include "phx_common.h"

namespace phx {

// NAIF ids of the bodies the force model knows (Earth is the only centre).
enum : int {
    NAIF_SSB = 0, NAIF_MERCURY_BARY = 1, NAIF_VENUS_BARY = 2, NAIF_EMB = 3,
    NAIF_MARS_BARY = 4, NAIF_JUPITER_BARY = 5, NAIF_SATURN_BARY = 6,
    NAIF_URANUS_BARY = 7, NAIF_NEPTUNE_BARY = 8, NAIF_PLUTO_BARY = 9,
    NAIF_SUN = 10, NAIF_MOON = 301, NAIF_EARTH = 399
};

// Fully (or un-) normalised spherical-harmonic field, triangular packed at
// n(n+1)/2+m, with the column-recursion factors interleaved as (a_nm, b_nm)
// and the sectoral factor c_m stored at the diagonal slot.
// Layout and factor definitions: interface/config/model/sphericalharmonics/
// Coefficients.h:69-107,480-535 (reference).
struct ShField {
    int degree = 0;
    int order = 0;
    bool normalized = true;
    double radius = 0.0;  // km
    double gm = 0.0;      // km^3/s^2
    std::vector<double> C, S;      // size (L+1)(L+2)/2
    std::vector<double> factors;   // size (L+1)(L+2), interleaved (a, b)
    size_t index(int n, int m) const { return size_t(n) * (n + 1) / 2 + m; }
};

// One daily CelesTrak SW-All record set, SoA by component
// (Nrlmsise00Atmosphere.h:18-34 of the reference: 13 components).
struct WeatherTable {
    enum { OBS_F107 = 0, OBS_F107A, ADJ_F107, ADJ_F107A, AVG_AP, AP0, NCOMP = 13 };
    double startMjdUtc = 0.0;
    int days = 0;
    std::vector<double> data;  // NCOMP * days
    double at(int comp, int record) const { return data[size_t(comp) * days + record]; }
};

// Daily EOP-All rows: xp, yp (arcsec), UT1-UTC, TAI-UTC (s).
struct EopTable {
    double firstMjdUtc = 0.0;
    int days = 0;
    std::vector<double> xp, yp, dut, dat;
};

struct NrlSettings {
    std::string spaceWeatherFile;
    double coverageStartMjdUtc = -1.0;
    double coverageEndMjdUtc = -1.0;
    double hCutoff = 1000.0;
    double densityScale = 1.0;
};

// One block of an exponential atmosphere: rho0 * exp(-(h - h0) / scaleHeight)
// for altitudes of its band, zero above hCutoff (rho0 in kg/km^3, lengths in km).
struct ExpSegment {
    double rho0 = 0.0, h0 = 0.0, scaleHeight = 1.0, hCutoff = 0.0;
};

struct BodySpec {
    std::string name;
    int naif = -1;
    bool pointGravity = true;
    bool shape = false;
    bool radiation = false;
    bool atmosphere = false;
    bool occulting = false;
    double gm = 0.0;          // km^3/s^2
    double radius = 0.0;      // km
    double soi = 1e20;        // km
    double flattening = 0.0;
    // gravity.shape
    std::string shFile;
    char shDelimiter = ',';
    int shHeaderLines = 0;
    int shDegree = -1;
    int shOrder = -1;
    // gravity.oblateness (J2 only; a shape field takes precedence)
    bool oblateness = false;
    // bodies.<name>.orientation: frame whose rotation the body's field uses
    // (empty: the unique loaded binary frame of the body, else its IAU model)
    std::string orientation;
    double j2 = 0.0;          // unnormalized J2 (constants.json "J2" or inline "j2")
    // atmosphere.nrlmsise00
    bool nrlmsise = false;
    NrlSettings nrl;
    // atmosphere.exponential: segments sorted ascending by h0
    std::vector<ExpSegment> expSegments;
};

struct Precision {
    bool mixed = false;   // FP32 terms for degrees above fp64Degree in orders above fp64Order
    int fp64Degree = 20;
    int fp64Order = 3;
};

struct ModelConfig {
    std::vector<BodySpec> bodies;
    std::string constantsFile;
    std::string ephemeris = "analytic";        // or "brie": Chebyshev tables from ephemerisFiles
    std::vector<std::string> ephemerisFiles;
    std::string orientation = "analytic";      // or "brot2": IERS2000 tables from orientationFiles
    std::vector<std::string> orientationFiles;
    bool preferSSB = true;
    // earthCorrections
    bool solidTides = false, oceanTides = false, relativity = false;
    bool zeroTide = true;
    std::string eopFile;
    // scene constants (constants.json header)
    double au = 149597870.7;
    double clight = 299792.458;
    double emrat = 81.30056822149722;

    int find(int naif) const {
        for (size_t i = 0; i < bodies.size(); ++i)
            if (bodies[i].naif == naif) return int(i);
        return -1;
    }
    bool tidesActive() const { return solidTides || oceanTides; }
    bool anyCorrections() const { return solidTides || oceanTides || relativity; }
};

// Run-level parameters from CONFIG.json ("run", "execution").
struct RunConfig {
    double durationSeconds = 0.0;
    double stepSeconds = 0.0;
    long long maxNumSteps = 0;
    int maxThreads = 1;
    Precision precision;
};

// A body other than the Earth whose field (spherical harmonics or J2) acts in
// its own rotating axes (physics/reduced/PrivateDimensional.h:527-559): the
// orientation unit of its resolved frame (naming/orientations/Resolver.h:45-92)
// and, for an IAU body unit, the nutation-precession unit of its barycenter.
constexpr int kMaxRotBodies = kMaxRotFields;
struct RotBody {
    int body = -1;           // index into ModelConfig::bodies
    int target = 0;          // rotation-data target (NAIF body or frame id)
    int unit = -1;           // index into BrotTable::bodies
    int bary = -1;           // type-1 units: barycenter unit (-1: none)
    bool sh = false;         // field from its coefficient file, else J2
    ShField field;
};

struct Environment {
    ModelConfig model;
    BrieTable brie;        // Chebyshev ephemeris (empty: analytic series)
    BrotTable brot;        // IERS2000 Earth orientation (empty: analytic model) and body units
    std::vector<RotBody> rot;
    ShField sh;            // Earth field (empty degree 0 when not configured)
    bool hasSh = false;
    WeatherTable weather;
    bool hasWeather = false;
    EopTable eop;
    bool hasEop = false;
};

// Lower-case copy (config keys match case-insensitively, as in the reference).
std::string lower(std::string s);

// Parse CONFIG.json. Relative data paths are resolved against the config's
// directory; every named data file must exist (driver.cpp:39-55).
void parseConfig(const json::Value& root, const std::string& configDir,
                 RunConfig& run, ModelConfig& model, bool requireRun);

// Load the data files named by the model configuration.
void loadEnvironment(Environment& env);

// Individual loaders (exposed for unit tests).
ShField loadShField(const std::string& path, char delimiter, int headerLines,
                    int degree, int order);
void shFactors(ShField& f);
WeatherTable loadSpaceWeather(const std::string& path, double requestedStart,
                              double requestedEnd);
EopTable loadEop(const std::string& path);
double calendarToMjd(int year, int month, int day);

}  // namespace phx
