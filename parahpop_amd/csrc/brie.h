// Chebyshev ephemeris tables in the reference's .brie format (CBOR), the
// position chains from a body to the Earth, and their evaluation. Shared by the
// host library (upload, coverage checks) and the CPU oracle.
//
// Reference behaviour followed (third_party/ unless noted):
//   file layouts 1 and 2                  brie/brie/util/cbor.cpp:200-300,
//                                         brie/brie/states/EphUnit.h:218-250,508-560
//   metadata (field-major, one column     brie/brie/states/metadata/RefEphUnit.h:13-44,
//   per body unit)                        brie/brie/states/metadata/EphUnit.h:50-100
//   body-unit data: radius, interval      brie/brie/core/BodyInterpolator.h:55-100
//   starts, coefficients [dim][deg][int]
//   interval and Chebyshev time           parm/parm/interpolate/detail/Interpolators.h:65-76,
//                                         parm/parm/interpolate/chebyshev/Interpolator.h:50-56
//   backward Clenshaw recurrence          parm/parm/interpolate/chebyshev/Interpolator.h:60-90
//   chains to the highest common center   brie/brie/core/ChainInterpolator.h:34-45,
//                                         brie/brie/states/metadata/RefEphUnit.h:401-430
#pragma once

#include <cmath>
#include <string>
#include <vector>

namespace phx {

// One body unit: `target` relative to `center` (NAIF ids), km, ICRF axes, ET seconds.
struct BrieUnit {
    int frame = 0, dtype = 0, target = 0, center = 0;
    int nIntervals = 0, pdeg = 0;
    long long offset = 0, size = 0;   // into BrieTable::data
    double start = 0.0, final = 0.0;
};

struct BrieTable {
    std::vector<BrieUnit> units;
    std::vector<double> data;
    bool empty() const { return units.empty(); }
};

// One step of a chain: add sign * position of unit `unit`.
struct BrieStep {
    int unit;
    double sign;
};

// IERS2000 Earth orientation unit of a .brot2 file (unit_type 4, model 0):
// dual-table Lagrange payload from `offset` (orientations/detail/LagrangeInterp.h:25-40):
//   X[nNut] Y[nNut] s[nNut] | xp[nErp] yp[nErp] dUT1[nErp] dX[nErp] dY[nErp] |
//   x0_nut h_nut x0_erp h_erp | ERP node keys[nErp]     (TT MJD2000 days)
struct BrotIers {
    int target = 0;
    int nNut = 0, nErp = 0;
    long long offset = 0;
    double start = 0.0, final = 0.0;   // envelope, ET seconds
};

// Body-orientation unit of a .brot2 file (orientations/metadata/Layout.h:19-60):
//   type 1  IAU body model (orientations/RefEphUnit.h:190-330): RA and Dec
//           polynomials in Julian centuries and the prime meridian polynomial
//           in days (degrees, coefficients in ascending powers), then the
//           nutation-precession amplitudes of RA, Dec and PM; p = {PDEG_RA,
//           PDEG_DEC, PDEG_PM, NUM_DIMS_RADII, BARYCENTER, NUM_NP_TERMS_RA,
//           NUM_NP_TERMS_DEC, NUM_NP_TERMS_PM}
//   type 2  nutation-precession angles of a barycenter: NUM_NP_ANGLES
//           polynomials of degree PDEG_NP in centuries (degrees); p = {PDEG_NP,
//           NUM_NP_ANGLES}
//   type 3  binary frame (core/Type2BodyUnit.h): Chebyshev 3-1-3 Euler angles
//           (radians) in the ephemeris body-unit layout; p = {INERT_FRAME,
//           INTERPOL_TYPE, NINTERVALS, PDEG}
struct BrotBody {
    int type = 0, target = 0;
    int p[8] = {};
    long long offset = 0, size = 0;   // into BrotTable::data
    double start = 0.0, final = 0.0;  // envelope, ET seconds
};

struct BrotTable {
    std::vector<BrotIers> iers;
    std::vector<BrotBody> bodies;
    std::vector<double> data;
    bool empty() const { return iers.empty(); }
    // the first unit of `target` (brie resolves queries through the first
    // matching row), -1 when none
    int findBody(int target) const {
        for (size_t i = 0; i < bodies.size(); ++i)
            if (bodies[i].target == target) return int(i);
        return -1;
    }
};

// Read one .brot2 file (layout 2) and append its IERS2000 (type 4, model 0) and
// body-orientation (types 1-3) units to `t`; throws when it holds neither.
void loadBrot2(const std::string& path, BrotTable& t);

// Read one .brie file (layout 1 or 2) and append its units to `t`.
void loadBrie(const std::string& path, BrieTable& t);

// The steps that give body `naif` relative to body `center`: the target's
// chain up to the highest common center (+), the center's chain up to it (-).
// Throws when either body has no unit.
std::vector<BrieStep> brieChain(const BrieTable& t, int naif, int center);

// [start, end] ET seconds that every unit of the chain covers.
void brieCoverage(const BrieTable& t, const std::vector<BrieStep>& chain, double& start, double& end);

// Position (km) of one unit at `et`, Clenshaw over the interval's coefficients.
inline void brieUnitPosition(const BrieTable& t, const BrieUnit& u, double et, double out[3]) {
    const double* d = t.data.data() + u.offset;
    const double r = d[0];
    long long idx = (long long)((et - d[1]) / (2.0 * r));
    if (idx < 0) idx = 0;
    if (idx > u.nIntervals - 1) idx = u.nIntervals - 1;
    const double s = (et - d[1 + idx]) / r - 1.0;
    const double s2 = 2.0 * s;
    const long long n = u.nIntervals;
    for (int k = 0; k < 3; ++k) {
        const double* c = d + 1 + n + (long long)k * n * (u.pdeg + 1) + idx;
        double w = 0.0, x = 0.0;
        for (int i = u.pdeg; i >= 1; --i) {
            const double tmp = c[i * n] - w;
            w = x;
            x = std::fma(s2, x, tmp);
        }
        out[k] = std::fma(s, x, c[0] - w);
    }
}

inline void brieChainPosition(const BrieTable& t, const std::vector<BrieStep>& chain, double et, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    for (const auto& st : chain) {
        double p[3];
        brieUnitPosition(t, t.units[size_t(st.unit)], et, p);
        for (int k = 0; k < 3; ++k) out[k] += st.sign * p[k];
    }
}

}  // namespace phx
