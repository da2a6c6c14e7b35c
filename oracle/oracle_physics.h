// ORACLE — test infrastructure only. CPU restatement of the reference host
// right-hand side, used as the parity checker and as the timed CPU baseline.
// Nothing in the product library includes or links this file.
//
// Each function cites the reference file:line it restates (paths relative to
// src/common/engine/ unless they start with third_party/).  Arithmetic order
// follows the reference host path so that a transcription slip shows up as a
// parity failure against the independently written HIP kernels.
//
// Functions without a reference counterpart (the analytic Sun/Moon/planet
// series and the analytic Earth orientation, which stand in for the .brie
// ephemeris and the .brot2 orientation files that the configurations do not
// provide) follow the published methods named at each function.
//
// What pins each part (tests/test_oracle_pins.py, tests/test_oracle_pins_env.py
// and the NRLMSISE-00 golden vectors): RKF78 tableau identities; closed-form
// Kepler propagation and the energy / angular-momentum invariants for the
// central-gravity driver; final step shortening; the SH sum against the
// analytic J2 formula (zonal field), finite differences of the potential and
// its own limit at the pole; the third body against its on-axis form and tidal
// limit; the shadow against disk-overlap areas, the annular 1 - b^2/a^2 and
// continuity; the relativity closed form; the analytic ephemeris against its
// published accuracy bands; the EOP time model against linear tables with a
// TAI-UTC step; the solid tide against IERS 2010 eqs. 6.6-6.8 and 6.13 and
// Tables 6.5a-c; the FES2004 sum against eq. 6.15 and its Doodson numbers; the
// Doodson arguments against the textbook mean longitudes; the pole tides
// against the IERS 2010 mean pole; the tide acceleration against finite
// differences of its potential. Parity unpinned: only the ULP-level behaviour
// of the device math library against glibc.
#pragma once

#include "flop_scope.h"

#include <cmath>
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

namespace oracle {

struct V3 { double x = 0, y = 0, z = 0; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// feta norms: squaredNorm = x*x + y*y + z*z; rCubedNorm = 1/(s*sqrt(s))
// (third_party/feta/feta/vector/expr/Expression.h:330-405)
inline double squaredNorm(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double norm(V3 a) { return std::sqrt(squaredNorm(a)); }
inline double rCubedNorm(V3 a) { const double s = squaredNorm(a); return 1.0 / (s * std::sqrt(s)); }

struct M3 { double m[9]; };  // row-major
inline V3 mul(const M3& R, V3 v) {
    return {R.m[0] * v.x + R.m[1] * v.y + R.m[2] * v.z,
            R.m[3] * v.x + R.m[4] * v.y + R.m[5] * v.z,
            R.m[6] * v.x + R.m[7] * v.y + R.m[8] * v.z};
}
inline V3 mulT(const M3& R, V3 v) {
    return {R.m[0] * v.x + R.m[3] * v.y + R.m[6] * v.z,
            R.m[1] * v.x + R.m[4] * v.y + R.m[7] * v.z,
            R.m[2] * v.x + R.m[5] * v.y + R.m[8] * v.z};
}
inline M3 matmul(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C.m[3 * i + j] = A.m[3 * i] * B.m[j] + A.m[3 * i + 1] * B.m[3 + j] + A.m[3 * i + 2] * B.m[6 + j];
    return C;
}

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kDeg = kPi / 180.0;
constexpr double kArcsec = kPi / (180.0 * 3600.0);

// ---------------------------------------------------------------------------
// RKF78 tableau — third_party/parm/parm/integrate/rk/coefs/tableau.h:19-51
// ---------------------------------------------------------------------------
struct Tableau {
    double c[12];
    double a[12][12];
    double b[13];
};
inline const Tableau& rkf78() {
    static const Tableau t = {
        {2. / 27, 1. / 9, 1. / 6, 5. / 12, 1. / 2, 5. / 6, 1. / 6, 2. / 3, 1. / 3, 1.0, 0, 1.0},
        {{2. / 27, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1. / 36, 1. / 12, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1. / 24, 0, 1. / 8, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {5. / 12, 0, -25. / 16, 25. / 16, 0, 0, 0, 0, 0, 0, 0, 0},
         {1. / 20, 0, 0, 1. / 4, 1. / 5, 0, 0, 0, 0, 0, 0, 0},
         {-25. / 108, 0, 0, 125. / 108, -65. / 27, 125. / 54, 0, 0, 0, 0, 0, 0},
         {31. / 300, 0, 0, 0, 61. / 225, -2. / 9, 13. / 900, 0, 0, 0, 0, 0},
         {2.0, 0, 0, -53. / 6, 704. / 45, -107. / 9, 67. / 90, 3.0, 0, 0, 0, 0},
         {-91. / 108, 0, 0, 23. / 108, -976. / 135, 311. / 54, -19. / 60, 17. / 6, -1. / 12, 0, 0, 0},
         {2383. / 4100, 0, 0, -341. / 164, 4496. / 1025, -301. / 82, 2133. / 4100, 45. / 82, 45. / 164, 18. / 41, 0, 0},
         {3. / 205, 0, 0, 0, 0, -6. / 41, -3. / 205, -3. / 41, 3. / 41, 6. / 41, 0, 0},
         {-1777. / 4100, 0, 0, -341. / 164, 4496. / 1025, -289. / 82, 2193. / 4100, 51. / 82, 33. / 164, 12. / 41, 0, 1.0}},
        {0, 0, 0, 0, 0, 34. / 105, 9. / 35, 9. / 35, 9. / 280, 9. / 280, 0, 41. / 840, 41. / 840}};
    return t;
}

// ---------------------------------------------------------------------------
// Time scales
// ---------------------------------------------------------------------------
// TimeConversion.h:30-40 (reference): ET = mjd*86400 - 43200; mjd = et/86400 + 0.5
inline double mjdToSpice(double mjd) { return mjd * 86400.0 + -43200.0; }
inline double spiceToMjd(double et) { return et * (1.0 / 86400.0) + 0.5; }

struct Eop {
    double first = 0.0;
    int days = 0;
    std::vector<double> xp, yp, dut, dat;
};

struct TimeData { double mjdUtc, ut1Utc, ttUtc, xpArcsec, ypArcsec; bool valid; };

// model/environment/earthcorrections/EarthCorrections.h:46-56
inline double tdbMinusTt(double mjdTt) {
    const double t = (mjdTt - 51544.5) / 36525;
    return 0.001657 * std::sin(628.3076 * t + 6.2401) + 0.000022 * std::sin(575.3385 * t + 4.2970) +
           0.000014 * std::sin(1256.6152 * t + 6.1969) + 0.000005 * std::sin(606.9777 * t + 4.0212) +
           0.000005 * std::sin(52.9691 * t + 0.4444) + 0.000002 * std::sin(21.3299 * t + 5.5431) +
           0.000010 * std::sin(628.3076 * t + 4.2490);
}

inline TimeData invalidTime() { const double q = std::nan(""); return {q, q, q, q, q, false}; }

// EarthCorrections.h:67-82
inline TimeData timeAtMjdUtc(const Eop& e, double mjdUtc) {
    if (e.days < 2 || !(mjdUtc >= e.first) || !(mjdUtc < e.first + double(e.days - 1))) return invalidTime();
    const int i = int(std::floor(mjdUtc - e.first));
    const double f = mjdUtc - std::floor(mjdUtc);
    const double xp = e.xp[i], yp = e.yp[i], dut = e.dut[i];
    return {mjdUtc, dut + f * (e.dut[i + 1] - dut), 32.184 + e.dat[i], xp + f * (e.xp[i + 1] - xp),
            yp + f * (e.yp[i + 1] - yp), true};
}

// EarthCorrections.h:87-105
inline TimeData timeAt(const Eop& e, double epochEt) {
    ORACLE_FLOP_SCOPE(kEpoch);
    if (e.days < 2 || !std::isfinite(epochEt)) return invalidTime();
    double ttEt = epochEt;
    for (int j = 0; j < 2; ++j) ttEt = epochEt - tdbMinusTt(51544.5 + ttEt / 86400);
    double utc = 51544.5 + (ttEt - 69.184) / 86400;
    for (int j = 0; j < 3; ++j) {
        double day = std::floor(utc - e.first);
        if (day < 0) day = 0;
        if (day > double(e.days - 1)) day = double(e.days - 1);
        const int i = int(day);
        utc = 51544.5 + (ttEt - 32.184 - e.dat[size_t(i)]) / 86400;
    }
    return timeAtMjdUtc(e, utc);
}

// ---------------------------------------------------------------------------
// Analytic Earth orientation (no reference counterpart: the configurations
// name no .brot2 file). ICRF -> ITRF: R = W(xp,yp)^T * R3(ERA) * C(X,Y,s)^T,
// IERS Conventions 2010 eqs. 5.1, 5.10, 5.15 with the IAU 2006 precession
// part of the CIP coordinates (Table 5.2a polynomial terms) and no nutation.
// ---------------------------------------------------------------------------
struct Orientation {
    M3 R;      // inertial -> Earth-fixed
    V3 omega;  // Earth angular velocity in the inertial frame, rad/s
};

constexpr double kEarthRate = 7.292115146706979e-5;  // 2*pi*1.00273781191135448/86400

inline M3 rotX(double a) { const double c = std::cos(a), s = std::sin(a); return {{1, 0, 0, 0, c, s, 0, -s, c}}; }
inline M3 rotY(double a) { const double c = std::cos(a), s = std::sin(a); return {{c, 0, -s, 0, 1, 0, s, 0, c}}; }
inline M3 rotZ(double a) { const double c = std::cos(a), s = std::sin(a); return {{c, s, 0, -s, c, 0, 0, 0, 1}}; }

// Earth rotation angle, IERS 2010 eq. 5.15 with the same MJD split as the
// reference's gmst06 helper (accelerations/hpop_compat/HpopCorrections.h:94-98).
inline double earthRotationAngle(double mjdUt1) {
    const double eraT = mjdUt1 + (2400000.5 - 2451545.0);
    const double f = std::fmod(mjdUt1, 1.0) + std::fmod(2400000.5, 1.0);
    double a = std::fmod(kTwoPi * (f + 0.7790572732640 + 0.00273781191135448 * eraT), kTwoPi);
    return a < 0 ? a + kTwoPi : a;
}

// Time arguments used by the orientation: UT1 and TT from the EOP model when
// it is loaded, else UTC = TT - 69.184 s, UT1 = UTC and zero polar motion.
inline void orientationTimes(const Eop* eop, double et, double& mjdUt1, double& tTT, double& xp, double& yp) {
    if (eop && eop->days >= 2) {
        const TimeData t = timeAt(*eop, et);
        mjdUt1 = t.mjdUtc + t.ut1Utc / 86400.0;
        tTT = (t.mjdUtc + t.ttUtc / 86400.0 - 51544.5) / 36525.0;
        xp = t.xpArcsec;
        yp = t.ypArcsec;
        return;
    }
    double ttEt = et;
    for (int j = 0; j < 2; ++j) ttEt = et - tdbMinusTt(51544.5 + ttEt / 86400);
    mjdUt1 = 51544.5 + (ttEt - 69.184) / 86400.0;
    tTT = ttEt / (86400.0 * 36525.0);
    xp = yp = 0.0;
}

inline Orientation earthOrientation(const Eop* eop, double et) {
    double mjdUt1, t, xp, yp;
    orientationTimes(eop, et, mjdUt1, t, xp, yp);
    const double X = (-0.016617 + t * (2004.191898 + t * (-0.4297829 + t * (-0.19861834)))) * kArcsec;
    const double Y = (-0.006951 + t * (-0.025896 + t * (-22.4072747 + t * 0.00190059))) * kArcsec;
    const double s = -0.5 * X * Y + (0.000094 + 0.00380865 * t) * kArcsec;
    const double a = 1.0 / (1.0 + std::sqrt(1.0 - X * X - Y * Y));
    const M3 Cb = {{1 - a * X * X, -a * X * Y, -X, -a * X * Y, 1 - a * Y * Y, -Y, X, Y, 1 - a * (X * X + Y * Y)}};
    // GCRS -> CIRS = R3(-s) R3(-E) R2(d) R3(E) (the transpose of eq. 5.10's Q)
    const M3 C = matmul(rotZ(-s), Cb);
    const M3 E = matmul(rotZ(earthRotationAngle(mjdUt1)), C);
    const double sp = -47e-6 * t * kArcsec;
    // TIRS -> ITRS = R1(-yp) R2(-xp) R3(s') (the transpose of eq. 5.3's W)
    const M3 W = matmul(rotX(-yp * kArcsec), matmul(rotY(-xp * kArcsec), rotZ(sp)));
    Orientation o;
    o.R = matmul(W, E);
    o.omega = mulT(o.R, V3{0.0, 0.0, kEarthRate});
    return o;
}

// ---------------------------------------------------------------------------
// IERS2000 Earth orientation from a .brot2 unit —
// brie/orientations/detail/IERS2000Model.h:120-300,380-400,505-560 (TT model
// time, prelude, R = PM ER PN, omega = vee(Rdot^T R)) and
// brie/orientations/detail/LagrangeInterp.h:25-150 with
// parm/interpolate/lagrange/Interpolator.h:45-95,416-432,815-835 (six-node
// uniform nutation stencil, four-node ERP stencil over the node keys).
// payload: X,Y,s [n1 each] | xp,yp,dUT1,dX,dY [n2 each] | x0,h,x0e,he | keys[n2]
// ---------------------------------------------------------------------------
// value and d/dq of the Lagrange polynomial through nodes (q_m, y_m), m < np
inline void lagrange(int np, const double* nodes, const double* y, int stride, int nch, double q, double* val, double* der) {
    for (int c = 0; c < nch; ++c) val[c] = der[c] = 0.0;
    for (int j = 0; j < np; ++j) {
        double prod = 1.0, dprod = 0.0, denom = 1.0;
        for (int m = 0; m < np; ++m) {
            if (m == j) continue;
            const double f = q - nodes[m];
            dprod = std::fma(dprod, f, prod);
            prod *= f;
            denom *= nodes[j] - nodes[m];
        }
        const double wv = prod * (1.0 / denom), wd = dprod * (1.0 / denom);
        for (int c = 0; c < nch; ++c) {
            val[c] = std::fma(wv, y[c * stride + j], val[c]);
            der[c] = std::fma(wd, y[c * stride + j], der[c]);
        }
    }
}

inline Orientation iers2000Orientation(const double* d, int n1, int n2, double et) {
    const double day = 86400.0;
    const double tdbMjd = et / day + 0.5;
    const double g = kDeg * (357.53 + 0.9856003 * (tdbMjd - 0.5));
    const double corr = (0.001658 * std::sin(g) + 0.000014 * std::sin(2.0 * g)) / day;
    double frac = std::fmod(et, day) / day + 0.5;
    frac -= std::floor(frac);
    const double ttFrac = frac - corr, tt = tdbMjd - corr;
    const double* k = d + 3 * n1 + 5 * n2;
    const double x0 = k[0], h = k[1];
    long lo = (long)std::floor((tt - x0) / h) - 2;
    lo = std::max(0L, std::min(lo, (long)n1 - 6));
    const double localNodes[6] = {0, 1, 2, 3, 4, 5};
    double nut[3], nutd[3];
    lagrange(6, localNodes, d + lo, n1, 3, (tt - (x0 + double(lo) * h)) / h, nut, nutd);
    const double* keys = k + 4;
    const long last = (long)(std::upper_bound(keys, keys + n2, tt) - keys) - 1;
    const long st = std::max(0L, std::min(last - 1, (long)n2 - 4));
    double erp[5], erpd[5];
    lagrange(4, keys + st, d + 3 * n1 + st, n2, 5, tt, erp, erpd);
    const double X = nut[0] + erp[3], Y = nut[1] + erp[4], s = nut[2];
    const double Xd = (nutd[0] / h + erpd[3]) / day, Yd = (nutd[1] / h + erpd[4]) / day, sd = nutd[2] / h / day;
    const double xp = erp[0], yp = erp[1], dut1 = erp[2];
    const double xpd = erpd[0] / day, ypd = erpd[1] / day, dut1d = erpd[2] / day;
    const double r2 = X * X + Y * Y, rr = std::sqrt(r2), fr = 1.0 / std::sqrt(1.0 - r2);
    const double e = r2 > 0 ? std::atan2(Y, X) : 0.0, dd = std::atan(rr * fr);
    const double edot = r2 > 0 ? (Yd * X - Xd * Y) / r2 : 0.0, ddot = (X * Xd + Y * Yd) * fr / rr;
    const double mm = -e - s, mmd = -edot - sd;
    double theta = kTwoPi * (std::fmod(ttFrac + dut1 - 0.5, 1.0) + 0.7790572732640 + 0.00273781191135448 * (tt + dut1 - 0.5));
    theta -= kTwoPi * std::floor(theta / kTwoPi);
    const double thetad = kTwoPi * ((1.0 / day + dut1d) + 0.00273781191135448 / day);
    const double sp = -6.24e-15 * (tt - 0.5), spd = -7.2205e-20;
    // PN = R3(mm) R2(dd) R3(e) and its derivative by the product rule
    const M3 PN = matmul(rotZ(mm), matmul(rotY(dd), rotZ(e)));
    auto dRotZ = [](double a, double ad) { const double c = std::cos(a), s = std::sin(a); return M3{{-s * ad, c * ad, 0, -c * ad, -s * ad, 0, 0, 0, 0}}; };
    auto dRotY = [](double a, double ad) { const double c = std::cos(a), s = std::sin(a); return M3{{-s * ad, 0, -c * ad, 0, 0, 0, c * ad, 0, -s * ad}}; };
    auto add = [](const M3& A, const M3& B) { M3 C; for (int i = 0; i < 9; ++i) C.m[i] = A.m[i] + B.m[i]; return C; };
    const M3 PNd = add(add(matmul(dRotZ(mm, mmd), matmul(rotY(dd), rotZ(e))), matmul(rotZ(mm), matmul(dRotY(dd, ddot), rotZ(e)))),
                       matmul(rotZ(mm), matmul(rotY(dd), dRotZ(e, edot))));
    const M3 ER = rotZ(theta), ERd = dRotZ(theta, thetad);
    const M3 PM = {{1.0, sp, xp, -sp, 1.0, -yp, -xp, yp, 1.0}};
    const M3 PMd = {{0.0, spd, xpd, -spd, 0.0, -ypd, -xpd, ypd, 0.0}};
    Orientation o;
    o.R = matmul(PM, matmul(ER, PN));
    const M3 Rd = add(add(matmul(PM, matmul(ERd, PN)), matmul(PMd, matmul(ER, PN))), matmul(PM, matmul(ER, PNd)));
    M3 M;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M.m[3 * i + j] = Rd.m[i] * o.R.m[j] + Rd.m[3 + i] * o.R.m[3 + j] + Rd.m[6 + i] * o.R.m[6 + j];
    o.omega = V3{0.5 * (M.m[7] - M.m[5]), 0.5 * (M.m[2] - M.m[6]), 0.5 * (M.m[3] - M.m[1])};
    return o;
}

// ---------------------------------------------------------------------------
// Analytic ephemeris (no reference counterpart: the configurations name
// "analytic mean-element series" in place of a .brie file).
//  * Sun and Moon: Montenbruck & Gill, Satellite Orbits (2000), sec. 3.3.2,
//    low-precision series referred to the mean ecliptic and equinox of J2000.
//  * Planetary system barycenters: E.M. Standish, "Keplerian Elements for
//    Approximate Positions of the Major Planets" (JPL), Table 1 (1800-2050),
//    heliocentric J2000 ecliptic elements; geocentric = heliocentric + Sun.
// Output: geocentric ICRF-aligned km. T = TDB Julian centuries from J2000.
// ---------------------------------------------------------------------------
constexpr double kObliquityJ2000 = 23.43929111 * kDeg;
constexpr double kAuKm = 149597870.7;

inline V3 eclipticToEquatorial(V3 e) {
    const double c = std::cos(kObliquityJ2000), s = std::sin(kObliquityJ2000);
    return {e.x, c * e.y - s * e.z, s * e.y + c * e.z};
}

inline double wrap360(double d) { d = std::fmod(d, 360.0); return d < 0 ? d + 360.0 : d; }

inline V3 sunGeocentric(double T) {
    const double M = wrap360(357.5256 + 35999.049 * T) * kDeg;
    const double lon = (282.9400 * kDeg) + M + (6892.0 * std::sin(M) + 72.0 * std::sin(2.0 * M)) * kArcsec;
    const double r = (149.619 - 2.499 * std::cos(M) - 0.021 * std::cos(2.0 * M)) * 1.0e6;
    return eclipticToEquatorial(V3{r * std::cos(lon), r * std::sin(lon), 0.0});
}

inline V3 moonGeocentric(double T) {
    const double L0 = wrap360(218.31617 + 481267.88088 * T - 1.3972 * T) * kDeg;
    const double l = wrap360(134.96292 + 477198.86753 * T) * kDeg;
    const double lp = wrap360(357.52543 + 35999.04944 * T) * kDeg;
    const double F = wrap360(93.27283 + 483202.01873 * T) * kDeg;
    const double D = wrap360(297.85027 + 445267.11135 * T) * kDeg;
    const double dlon = 22640 * std::sin(l) + 769 * std::sin(2 * l) - 4586 * std::sin(l - 2 * D) +
                        2370 * std::sin(2 * D) - 668 * std::sin(lp) - 412 * std::sin(2 * F) -
                        212 * std::sin(2 * l - 2 * D) - 206 * std::sin(l + lp - 2 * D) + 192 * std::sin(l + 2 * D) -
                        165 * std::sin(lp - 2 * D) + 148 * std::sin(l - lp) - 125 * std::sin(D) -
                        110 * std::sin(l + lp) - 55 * std::sin(2 * F - 2 * D);
    const double lon = L0 + dlon * kArcsec;
    const double lat = (18520 * std::sin(F + lon - L0 + (412 * std::sin(2 * F) + 541 * std::sin(lp)) * kArcsec) -
                        526 * std::sin(F - 2 * D) + 44 * std::sin(l + F - 2 * D) - 31 * std::sin(-l + F - 2 * D) -
                        25 * std::sin(-2 * l + F) - 23 * std::sin(lp + F - 2 * D) + 21 * std::sin(-l + F) +
                        11 * std::sin(-lp + F - 2 * D)) * kArcsec;
    const double r = 385000 - 20905 * std::cos(l) - 3699 * std::cos(2 * D - l) - 2956 * std::cos(2 * D) -
                     570 * std::cos(2 * l) + 246 * std::cos(2 * l - 2 * D) - 205 * std::cos(lp - 2 * D) -
                     171 * std::cos(l + 2 * D) - 152 * std::cos(l + lp - 2 * D);
    return eclipticToEquatorial(V3{r * std::cos(lon) * std::cos(lat), r * std::sin(lon) * std::cos(lat), r * std::sin(lat)});
}

// Standish Table 1: a[au], e, I[deg], L[deg], long.peri[deg], long.node[deg] and rates per century.
struct PlanetElements { int naif; double el[6]; double rate[6]; };
inline const PlanetElements* planetElements(int naif) {
    static const PlanetElements table[] = {
        {1, {0.38709927, 0.20563593, 7.00497902, 252.25032350, 77.45779628, 48.33076593},
         {0.00000037, 0.00001906, -0.00594749, 149472.67411175, 0.16047689, -0.12534081}},
        {2, {0.72333566, 0.00677672, 3.39467605, 181.97909950, 131.60246718, 76.67984255},
         {0.00000390, -0.00004107, -0.00078890, 58517.81538729, 0.00268329, -0.27769418}},
        {4, {1.52371034, 0.09339410, 1.84969142, -4.55343205, -23.94362959, 49.55953891},
         {0.00001847, 0.00007882, -0.00813131, 19140.30268499, 0.44441088, -0.29257343}},
        {5, {5.20288700, 0.04838624, 1.30439695, 34.39644051, 14.72847983, 100.47390909},
         {-0.00011607, -0.00013253, -0.00183714, 3034.74612775, 0.21252668, 0.20469106}},
        {6, {9.53667594, 0.05386179, 2.48599187, 49.95424423, 92.59887831, 113.66242448},
         {-0.00125060, -0.00050991, 0.00193609, 1222.49362201, -0.41897216, -0.28867794}},
        {7, {19.18916464, 0.04725744, 0.77263783, 313.23810451, 170.95427630, 74.01692503},
         {-0.00196176, -0.00004397, -0.00242939, 428.48202785, 0.40805281, 0.04240589}},
        {8, {30.06992276, 0.00859048, 1.77004347, -55.12002969, 44.96476227, 131.78422574},
         {0.00026291, 0.00005105, 0.00035372, 218.45945325, -0.32241464, -0.00508664}},
        {9, {39.48211675, 0.24882730, 17.14001206, 238.92903833, 224.06891629, 110.30393684},
         {-0.00031596, 0.00005170, 0.00004818, 145.20780515, -0.04062942, -0.01183482}},
    };
    for (const auto& p : table) if (p.naif == naif) return &p;
    return nullptr;
}

inline V3 planetHeliocentric(const PlanetElements& p, double T) {
    double el[6];
    for (int k = 0; k < 6; ++k) el[k] = p.el[k] + p.rate[k] * T;
    const double a = el[0] * kAuKm, e = el[1], I = el[2] * kDeg;
    const double varpi = el[4] * kDeg, Omega = el[5] * kDeg;
    const double w = varpi - Omega;
    double M = std::fmod((el[3] - el[4]) * kDeg, kTwoPi);
    if (M > kPi) M -= kTwoPi;
    if (M < -kPi) M += kTwoPi;
    double E = M + e * std::sin(M);
    for (int it = 0; it < 6; ++it) E -= (E - e * std::sin(E) - M) / (1.0 - e * std::cos(E));
    const double xp = a * (std::cos(E) - e), yp = a * std::sqrt(1.0 - e * e) * std::sin(E);
    const double cw = std::cos(w), sw = std::sin(w), cO = std::cos(Omega), sO = std::sin(Omega), cI = std::cos(I), sI = std::sin(I);
    const V3 ecl{(cw * cO - sw * sO * cI) * xp + (-sw * cO - cw * sO * cI) * yp,
                 (cw * sO + sw * cO * cI) * xp + (-sw * sO + cw * cO * cI) * yp,
                 (sw * sI) * xp + (cw * sI) * yp};
    return eclipticToEquatorial(ecl);
}

// Geocentric position of a body by NAIF id (Earth -> zero).
inline V3 bodyGeocentric(int naif, double et) {
    const double T = et / (86400.0 * 36525.0);
    if (naif == 399) return V3{};
    if (naif == 10) return sunGeocentric(T);
    if (naif == 301) return moonGeocentric(T);
    const PlanetElements* p = planetElements(naif);
    if (!p) return V3{std::nan(""), std::nan(""), std::nan("")};
    return planetHeliocentric(*p, T) + sunGeocentric(T);
}

// ---------------------------------------------------------------------------
// Point gravity — accelerations/PointGravity.h:16-21,76-81
// ---------------------------------------------------------------------------
inline V3 twoBody(V3 d, double gm) {
    const double rc = rCubedNorm(d);
    return {(-gm * d.x) * rc, (-gm * d.y) * rc, (-gm * d.z) * rc};
}
inline V3 thirdBodyNB(V3 sc, V3 body, double gm) { return twoBody(sc - body, gm) + twoBody(body, gm); }

// ---------------------------------------------------------------------------
// Spherical harmonics — accelerations/sphericalharmonics/Evaluation.h:111-159,
// 282-296; DegreeLoop.h:51-132; LegendreRecursor.h:16-71,135-199;
// PseudoPhasor.h:26-109; polar limit Coefficients.h:123-132.
// ---------------------------------------------------------------------------
struct ShTables {
    int degree = 0, order = 0;
    bool normalized = true;
    double radius = 0, gm = 0;
    const double* C = nullptr;
    const double* S = nullptr;
    const double* factors = nullptr;  // interleaved (a, b); c_m at the diagonal
};

inline V3 shBodyFixed(V3 p, const ShTables& f) {
    const double rn = norm(p);
    const double sinPhi = p.z / rn;
    const double rho = std::sqrt(p.x * p.x + p.y * p.y);
    const double cosPhi = rho / rn;
    const double sinL = rho > 1e-14 ? p.y / rho : 0.0;
    const double cosL = rho > 1e-14 ? p.x / rho : 1.0;
    double sinML = 0.0, cosML = 1.0;
    const double alpha = f.radius / rn;
    double acc0 = 0, acc1 = 0, acc2 = 0;
    double alpham = 1.0;
    double Pmm = 1.0, dPmm = 0.0;
    for (int m = 0; m <= f.order; ++m) {
        double P = Pmm, dP = dPmm, Pn1 = 0.0, dPn1 = 0.0;
        double r0 = 0, r1 = 0, r2 = 0;
        double alphan = alpham;
        for (int n = m; n <= f.degree; ++n) {
            double divCos = (m == 0 || cosPhi == 0.0) ? 0.0 : P / cosPhi;
            const size_t idx = size_t(n) * (n + 1) / 2 + m;
            const double c = f.C[idx];
            const double s = m > 0 ? f.S[idx] : 0.0;
            const double phC = c * cosML + s * sinML;
            const double phS = c * sinML - s * cosML;
            if (m == 1 && cosPhi == 0.0) {
                const double sign = sinPhi > 0.0 ? 1.0 : ((n & 1) ? 1.0 : -1.0);
                double v = 0.5 * double(n) * (n + 1.0);
                if (f.normalized) v *= std::sqrt(2.0 * (2.0 * n + 1.0) / (double(n) * (n + 1.0)));
                divCos = sign * v;
            }
            r0 -= (double(n + 1) * alphan * P * phC);
            r1 += (alphan * dP * phC);
            r2 -= (double(m) * alphan * divCos * phS);
            if (n < f.degree) {
                const size_t fi = 2 * (size_t(n + 1) * (n + 2) / 2 + m);
                const double a = f.factors[fi], b = f.factors[fi + 1];
                const double Pn2 = Pn1;
                Pn1 = P;
                P = a * sinPhi * Pn1 - b * Pn2;
                const double dPn2 = dPn1;
                dPn1 = dP;
                dP = a * (sinPhi * dPn1 + cosPhi * Pn1) - b * dPn2;
                alphan *= alpha;
            }
        }
        acc0 += r0;
        acc1 += r1;
        acc2 += r2;
        if (m < f.order) {
            const double nS = cosL * sinML + sinL * cosML;
            const double nC = cosL * cosML - sinL * sinML;
            sinML = nS;
            cosML = nC;
            const double cm = f.factors[2 * (size_t(m + 1) * (m + 2) / 2 + m + 1)];
            const double newP = cm * cosPhi * Pmm;
            const double newdP = cm * (cosPhi * dPmm - sinPhi * Pmm);
            Pmm = newP;
            dPmm = newdP;
            alpham *= alpha;
        }
    }
    const double g = f.gm / (rn * rn);
    acc0 *= g;
    acc1 *= g;
    acc2 *= g;
    return {cosPhi * cosL * acc0 - sinPhi * cosL * acc1 - sinL * acc2,
            cosPhi * sinL * acc0 - sinPhi * sinL * acc1 + cosL * acc2,
            sinPhi * acc0 + cosPhi * acc1};
}

// SphericalHarmonics.h:28-41: rotate in, evaluate, rotate back.
inline V3 sphericalHarmonics(V3 pos, const M3& R, const ShTables& f) {
    return mulT(R, shBodyFixed(mul(R, pos), f));
}

// ---------------------------------------------------------------------------
// J2-only gravity — accelerations/J2.h:17-92 (Wakker eq. 20.6): skipped beyond
// the body's sphere of influence, evaluated in body-fixed axes.
// ---------------------------------------------------------------------------
inline V3 j2BodyFixed(V3 p, double gm, double j2, double radius) {
    const double invr = 1.0 / std::sqrt(squaredNorm(p));
    const double invr2 = invr * invr;
    const double invr5 = invr2 * invr2 * invr;
    const double v = 1.0 - 5.0 * (p.z * p.z) * invr2;
    const double c = -1.5 * gm * j2 * invr5 * radius * radius;
    return {c * p.x * v, c * p.y * v, c * p.z * (v + 2.0)};
}

// ---------------------------------------------------------------------------
// Axes of a body other than the Earth from its .brot2 orientation unit:
// orientations/RefEphUnit.h:190-420 (IAU body model with the nutation-precession
// sums of core/Polynomial.h:30-115; binary frame Euler angles), frames/Frames.h:
// 40-130,250-270 (radians, ECLIPJ2000 bias) and frames/Rotation.h:40-210 (the
// passive 3-1-3 rotation as the product of half-angle quaternions q3 q2 q1).
// ---------------------------------------------------------------------------
struct Quat { double w, x, y, z; };
inline Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// passive (frame) rotation by theta about principal axis `axis` (1 x, 3 z)
inline Quat passiveAxis(int axis, double theta) {
    const double s = std::sin(0.5 * theta), c = std::cos(0.5 * theta);
    return axis == 1 ? Quat{c, -s, 0.0, 0.0} : Quat{c, 0.0, 0.0, -s};
}
// the matrix of v -> q v q^-1
inline M3 quatMatrix(const Quat& q) {
    const double w = q.w, x = q.x, y = q.y, z = q.z;
    return {{1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
             2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
             2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)}};
}

// ICRF -> body axes from RA, Dec, W (radians)
inline M3 raDecPmAxes(double ra, double dec, double pm, bool eclipticBase) {
    const double halfPi = M_PI / 2;
    const Quat q1 = passiveAxis(3, halfPi + ra), q2 = passiveAxis(1, halfPi - dec), q3 = passiveAxis(3, pm);
    Quat q = qmul(qmul(q3, q2), q1);
    if (eclipticBase) q = qmul(q, passiveAxis(1, 84381.448 / 3600.0 * (M_PI / 180.0)));
    return quatMatrix(q);
}

// IAU body unit: d = RA[degRa+1] Dec[degDec+1] PM[degPm+1] ampRA[nRa] ampDec[nDec]
// ampPM[nPm] (ascending powers; RA, Dec in Julian centuries, PM in days past
// J2000 TDB; degrees); ang = the barycenter's angle polynomials (nullptr: none)
inline M3 iauBodyAxes(const double* d, const int deg[3], const int np[3], const double* ang, int angDeg, double et) {
    const double tc = et / 3155760000.0, td = et / 86400.0;
    auto poly = [](const double* c, int n, double t) {
        double r = 0.0;
        for (int i = n - 1; i >= 0; --i) r = std::fma(r, t, c[i]);
        return r;
    };
    auto trig = [&](const double* amp, int n, bool useSin) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) {
            double v = 0.0, tp = 1.0;
            for (int j = 0; j <= angDeg; ++j) {
                v = std::fma(ang[i * (angDeg + 1) + j], tp, v);
                tp *= tc;
            }
            r = std::fma(amp[i], useSin ? std::sin(v * (M_PI / 180.0)) : std::cos(v * (M_PI / 180.0)), r);
        }
        return r;
    };
    double ra = poly(d, deg[0] + 1, tc);
    double dec = poly(d + deg[0] + 1, deg[1] + 1, tc);
    double pm = poly(d + deg[0] + 1 + deg[1] + 1, deg[2] + 1, td);
    if (ang) {
        const double* amp = d + (deg[0] + 1) + (deg[1] + 1) + (deg[2] + 1);
        ra += trig(amp, np[0], true);
        dec += trig(amp + np[0], np[1], false);
        pm += trig(amp + np[0] + np[1], np[2], true);
    }
    const double k = M_PI / 180.0;
    return raDecPmAxes(std::fmod(ra, 360.0) * k, std::fmod(dec, 360.0) * k, std::fmod(pm, 360.0) * k, false);
}

// Binary frame unit: Chebyshev Euler angles (phi, theta, psi) in the ephemeris
// body-unit layout (radius, interval starts, coefficients [angle][degree][interval]),
// evaluated here by the forward recurrence T_{k+1} = 2 s T_k - T_{k-1}.
inline M3 binaryFrameAxes(const double* d, int nIntervals, int pdeg, int inertFrame, double et) {
    const double r = d[0];
    long long idx = (long long)std::floor((et - d[1]) / (2.0 * r));
    if (idx < 0) idx = 0;
    if (idx > nIntervals - 1) idx = nIntervals - 1;
    const double s = (et - d[1 + idx]) / r - 1.0;
    double e[3];
    for (int k = 0; k < 3; ++k) {
        double t0 = 1.0, t1 = s, v = 0.0;
        for (int i = 0; i <= pdeg; ++i) {
            const double c = d[1 + nIntervals + (long long)k * nIntervals * (pdeg + 1) + (long long)i * nIntervals + idx];
            const double T = i == 0 ? 1.0 : (i == 1 ? s : 2.0 * s * t1 - t0);
            if (i >= 2) { t0 = t1; t1 = T; }
            v += c * T;
        }
        e[k] = v;
    }
    const double halfPi = M_PI / 2;
    const double ra = e[0] - halfPi, dec = halfPi - e[1], pm = std::fmod(e[2], 2.0 * M_PI);
    return raDecPmAxes(ra, dec, pm, inertFrame == 17);
}

inline V3 j2Gravity(V3 pos, const M3& R, double gm, double j2, double radius, double soi) {
    if (norm(pos) > soi) return V3{};
    return mulT(R, j2BodyFixed(mul(R, pos), gm, j2, radius));
}

// ---------------------------------------------------------------------------
// Exponential atmosphere — environment/atmosphere/Geodetic.h:40-72 (geodetic
// altitude by fixed-point iteration on the polar coordinate; ||r|| - R_eq when
// the flattening is 0) and PiecewiseExponentialAtmosphere.h:100-170 (band =
// number of bands s >= 1 with h0_s <= h; zero above the band's cutoff).
// The iteration is capped at kGeodeticMaxIter (it converges in 2-4 steps).
// ---------------------------------------------------------------------------
constexpr int kGeodeticMaxIter = 32;

inline double geodeticAltitude(V3 r, V3 pole, double Req, double f) {
    if (f == 0.0) return norm(r) - Req;
    const double r2 = dot(r, r);
    const double z = dot(r, pole);
    const double xy = r2 - z * z;
    const double e2 = (2.0 - f) * f;
    double d0 = 0.0, d = 1e10, sfq = 0.0, zd = z;
    for (int it = 0; it < kGeodeticMaxIter && std::fabs(d - d0) > 1e-9; ++it) {
        d0 = d;
        const double z2 = zd * zd;
        sfq = z2 / (xy + z2);
        d = Req * e2 * std::sqrt(sfq / (1.0 - e2 * sfq));
        zd = z + std::copysign(d, z);
    }
    return std::sqrt(xy + zd * zd) - Req / std::sqrt(1.0 - e2 * sfq);
}

// S: a band with fields rho0, h0, scaleHeight, hCutoff
template <class S>
inline int expBand(const S* b, int n, double h) {
    int band = 0;
    for (int s = 1; s < n; ++s) band += b[s].h0 <= h ? 1 : 0;
    return band;
}

template <class S>
inline double expDensity(const S* b, int n, double h) {
    const S& s = b[expBand(b, n, h)];
    if (h > s.hCutoff) return 0.0;
    return s.rho0 * std::exp(-(h - s.h0) / s.scaleHeight);
}

// ---------------------------------------------------------------------------
// SRP and dual-cone shadow — accelerations/SRP.h:18,62-69;
// environment/eclipse/Shadow.h:84-142
// ---------------------------------------------------------------------------
inline V3 srp(V3 d, double AMS, double AU) {
    double factor = AU;
    factor *= factor * 4.56e-3 * AMS;
    const double k = factor * rCubedNorm(d);
    return {k * d.x, k * d.y, k * d.z};
}

inline double clampd(double x, double lo, double hi) { return std::fmin(std::fmax(x, lo), hi); }

inline double shadowFactor(V3 rS, V3 rB, double RSun, double RBody) {
    const double s2 = squaredNorm(rS), b2 = squaredNorm(rB), sb = dot(rS, rB);
    const double rInvS = 1.0 / std::sqrt(s2), rInvB = 1.0 / std::sqrt(b2);
    const bool inside = b2 <= RBody * RBody;
    const double a = std::asin(clampd(RSun * rInvS, 0.0, 1.0));
    const double b = std::asin(clampd(RBody * rInvB, 0.0, 1.0));
    const double c = std::acos(clampd(sb * rInvS * rInvB, -1.0, 1.0));
    if (inside) return 0.0;
    if (c >= a + b) return 1.0;
    if (c <= std::fabs(b - a)) {
        if (b >= a) return 0.0;
        return 1.0 - (b * b) / (a * a);
    }
    const double a2 = a * a, b2_ = b * b, c2 = c * c;
    const double x = (c2 + a2 - b2_) / (2.0 * c);
    const double y = std::sqrt(std::fmax(a2 - x * x, 0.0));
    const double alpha = std::acos(clampd(x / a, -1.0, 1.0));
    const double beta = std::acos(clampd((c - x) / b, -1.0, 1.0));
    const double occluded = a2 * alpha + b2_ * beta - c * y;
    return 1.0 - occluded / (a2 * 3.14159265358979323846);
}

// ---------------------------------------------------------------------------
// Earth corrections — accelerations/hpop_compat/HpopCorrections.h:19-392,
// accelerations/EarthCorrections.h:13-103
// ---------------------------------------------------------------------------
struct TideConstants {
    double muEarth = 398600.4415e9;
    double radius = 6378.1363e3;
    double muSun = 132712440041.279419e9;
    double muMoon = 398600.4415e9 / 81.3005682214972154;
    double c = 299792457.999999984;
    bool zeroTide = true;
};

struct FreqTerm { int order; int mult[5]; double inPhase, outPhase; };
struct OceanWave { int doodson[6]; int begin, end; };
struct OceanTerm { int n, m; double cp, sp, cm, sm; };

struct TideTables {
    std::vector<FreqTerm> freq[3];
    std::vector<OceanWave> waves;
    std::vector<OceanTerm> terms;
};

struct CS7 { double C[7][7]; double S[7][7]; };

inline double positiveMod(double x, double period) {
    const double a = std::fmod(x, period);
    return a < 0.0 ? a + period : a;
}

// HpopCorrections.h:94-103
inline double gmst06(double mjdUt1, double mjdTt) {
    const double eraT = mjdUt1 + (2400000.5 - 2451545.0);
    const double f = std::fmod(mjdUt1, 1.0) + std::fmod(2400000.5, 1.0);
    const double era = positiveMod(kTwoPi * (f + 0.7790572732640 + 0.00273781191135448 * eraT), kTwoPi);
    const double t = ((2400000.5 - 2451545.0) + mjdTt) / 36525.0;
    return positiveMod(era + (0.014506 + (4612.156534 + (1.3915817 + (-0.00000044 + (-0.000029956 - 0.0000000368 * t) * t) * t) * t) * t) * kArcsec, kTwoPi);
}

// HpopCorrections.h:105-119
inline void legendreValues(int nmax, double phi, double p[7][7]) {
    for (int n = 0; n < 7; ++n) for (int m = 0; m < 7; ++m) p[n][m] = 0.0;
    const double s = std::sin(phi), c = std::cos(phi);
    p[0][0] = 1.0;
    if (nmax == 0) return;
    p[1][1] = std::sqrt(3.0) * c;
    for (int n = 2; n <= nmax; ++n) p[n][n] = std::sqrt((2.0 * n + 1) / (2.0 * n)) * c * p[n - 1][n - 1];
    for (int n = 1; n <= nmax; ++n) p[n][n - 1] = std::sqrt(2.0 * n + 1) * s * p[n - 1][n - 1];
    for (int m = 0; m <= nmax; ++m)
        for (int n = m + 2; n <= nmax; ++n)
            p[n][m] = std::sqrt((2.0 * n + 1) / ((n - m) * double(n + m))) *
                      (std::sqrt(2.0 * n - 1) * s * p[n - 1][m] - std::sqrt(((n + m - 1) * double(n - m - 1)) / (2.0 * n - 3)) * p[n - 2][m]);
}

// HpopCorrections.h:121-145
inline void legendre(int nmax, double phi, double p[7][7], double dp[7][7]) {
    for (int n = 0; n < 7; ++n) for (int m = 0; m < 7; ++m) { p[n][m] = 0.0; dp[n][m] = 0.0; }
    const double s = std::sin(phi), c = std::cos(phi);
    p[0][0] = 1.0;
    if (nmax == 0) return;
    p[1][1] = std::sqrt(3.0) * c;
    dp[1][1] = -std::sqrt(3.0) * s;
    for (int n = 2; n <= nmax; ++n) {
        const double k = std::sqrt((2.0 * n + 1) / (2.0 * n));
        p[n][n] = k * c * p[n - 1][n - 1];
        dp[n][n] = k * (c * dp[n - 1][n - 1] - s * p[n - 1][n - 1]);
    }
    for (int n = 1; n <= nmax; ++n) {
        const double k = std::sqrt(2.0 * n + 1);
        p[n][n - 1] = k * s * p[n - 1][n - 1];
        dp[n][n - 1] = k * (c * p[n - 1][n - 1] + s * dp[n - 1][n - 1]);
    }
    for (int m = 0; m <= nmax; ++m)
        for (int n = m + 2; n <= nmax; ++n) {
            const double k = std::sqrt((2.0 * n + 1) / ((n - m) * double(n + m)));
            const double a = std::sqrt(2.0 * n - 1);
            const double b = std::sqrt(((n + m - 1) * double(n - m - 1)) / (2.0 * n - 3));
            p[n][m] = k * (a * s * p[n - 1][m] - b * p[n - 2][m]);
            dp[n][m] = k * (a * s * dp[n - 1][m] + a * c * p[n - 1][m] - b * dp[n - 2][m]);
        }
}

struct TideInput {
    double r[3], v[3], sun[3], moon[3];
    double E[9];
    double mjdUtc, ut1Utc, ttUtc, xpArcsec, ypArcsec;
};

inline void rot(const double E[9], const double x[3], double out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = E[3 * i] * x[0] + E[3 * i + 1] * x[1] + E[3 * i + 2] * x[2];
}
inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// HpopCorrections.h:175-184
inline void tideArguments(const TideInput& in, double args[5], double& thetaG) {
    const double mjdTt = in.mjdUtc + in.ttUtc / 86400.0;
    const double t = (mjdTt - 51544.5) / 36525.0, t2 = t * t, t3 = t2 * t, t4 = t3 * t;
    args[0] = positiveMod(485868.249036 + 1717915923.2178 * t + 31.8792 * t2 + 0.051635 * t3 - 0.00024470 * t4, 1296000.0) * kArcsec;
    args[1] = positiveMod(1287104.79305 + 129596581.0481 * t - 0.5532 * t2 + 0.000136 * t3 - 0.00001149 * t4, 1296000.0) * kArcsec;
    args[2] = positiveMod(335779.526232 + 1739527262.8478 * t - 12.7512 * t2 - 0.001037 * t3 + 0.00000417 * t4, 1296000.0) * kArcsec;
    args[3] = positiveMod(1072260.70369 + 1602961601.2090 * t - 6.3706 * t2 + 0.006593 * t3 - 0.00003169 * t4, 1296000.0) * kArcsec;
    args[4] = positiveMod(450160.398036 - 6962890.5431 * t + 7.4722 * t2 + 0.007702 * t3 - 0.00005939 * t4, 1296000.0) * kArcsec;
    thetaG = gmst06(in.mjdUtc + in.ut1Utc / 86400.0, mjdTt);
}

// HpopCorrections.h:187-199
inline void poleWobble(const TideInput& in, double& m1, double& m2) {
    const double t = (in.mjdUtc + in.ttUtc / 86400.0 - 51544.5) / 365.25;
    double x, y;
    if (t < 10.0) {
        x = 55.974 + t * (1.8243 + t * (0.18413 + t * 0.007024));
        y = 346.346 + t * (1.7896 + t * (-0.10729 - t * 0.000908));
    } else {
        x = 23.513 + 7.6141 * t;
        y = 358.891 - 0.6287 * t;
    }
    m1 = in.xpArcsec - 1e-3 * x;
    m2 = -(in.ypArcsec - 1e-3 * y);
}

inline void clearCS(CS7& o) { for (int n = 0; n < 7; ++n) for (int m = 0; m < 7; ++m) o.C[n][m] = o.S[n][m] = 0.0; }

// HpopCorrections.h:148-169,201-264,326-329
inline void solidCoefficients(const TideInput& in, const TideConstants& k, const TideTables& tt, CS7& out) {
    ORACLE_FLOP_SCOPE(kEpoch);
    double moon[3], sun[3];
    rot(in.E, in.moon, moon);
    rot(in.E, in.sun, sun);
    const double moonLon = std::atan2(moon[1], moon[0]);
    const double sunLon = std::atan2(sun[1], sun[0]);
    const double moonRatio = k.radius / std::sqrt(dot3(moon, moon));
    const double sunRatio = k.radius / std::sqrt(dot3(sun, sun));
    double moonP[7][7], sunP[7][7];
    legendreValues(3, std::atan2(moon[2], std::hypot(moon[0], moon[1])), moonP);
    legendreValues(3, std::atan2(sun[2], std::hypot(sun[0], sun[1])), sunP);
    clearCS(out);
    const double moonFactor = k.muMoon / k.muEarth * std::pow(moonRatio, 3.0);
    const double sunFactor = k.muSun / k.muEarth * std::pow(sunRatio, 3.0);
    double qC[3], qS[3];
    for (int m = 0; m <= 2; ++m) {
        qC[m] = moonFactor * moonP[2][m] * std::cos(m * moonLon) + sunFactor * sunP[2][m] * std::cos(m * sunLon);
        qS[m] = moonFactor * moonP[2][m] * std::sin(m * moonLon) + sunFactor * sunP[2][m] * std::sin(m * sunLon);
    }
    out.C[2][0] = 0.30190 / 5.0 * qC[0];
    out.C[2][1] = 0.29830 / 5.0 * qC[1] - 0.00144 / 5.0 * qS[1];
    out.S[2][1] = 0.00144 / 5.0 * qC[1] + 0.29830 / 5.0 * qS[1];
    out.C[2][2] = 0.30102 / 5.0 * qC[2] - 0.00130 / 5.0 * qS[2];
    out.S[2][2] = 0.00130 / 5.0 * qC[2] + 0.30102 / 5.0 * qS[2];
    out.C[4][0] = -0.00089 / 5.0 * qC[0];
    out.C[4][1] = -0.00080 / 5.0 * qC[1];
    out.S[4][1] = -0.00080 / 5.0 * qS[1];
    out.C[4][2] = -0.00057 / 5.0 * qC[2];
    out.S[4][2] = -0.00057 / 5.0 * qS[2];
    const double moon3 = k.muMoon / k.muEarth * std::pow(moonRatio, 4.0);
    const double sun3 = k.muSun / k.muEarth * std::pow(sunRatio, 4.0);
    for (int m = 0; m <= 3; ++m) {
        const double factor = (m == 3 ? 0.094 : 0.093) / 7.0;
        out.C[3][m] = factor * (moon3 * moonP[3][m] * std::cos(m * moonLon) + sun3 * sunP[3][m] * std::cos(m * sunLon));
        if (m) out.S[3][m] = factor * (moon3 * moonP[3][m] * std::sin(m * moonLon) + sun3 * sunP[3][m] * std::sin(m * sunLon));
    }
    double args[5], thetaG;
    tideArguments(in, args, thetaG);
    for (int order = 0; order <= 2; ++order) {
        double dC = 0.0, dS = 0.0;
        for (const FreqTerm& term : tt.freq[order]) {
            double arg = 0.0;
            for (int j = 0; j < 5; ++j) arg += term.mult[j] * args[j];
            const double phase = order * (thetaG + kPi) - arg;
            const double c = std::cos(phase), s = std::sin(phase);
            if (order == 0) dC += 1e-12 * (term.inPhase * c - term.outPhase * s);
            else if (order == 1) {
                dC += 1e-12 * (term.inPhase * s + term.outPhase * c);
                dS += 1e-12 * (term.inPhase * c - term.outPhase * s);
            } else {
                dC += 1e-12 * term.inPhase * c;
                dS -= 1e-12 * term.inPhase * s;
            }
        }
        out.C[2][order] += dC;
        out.S[2][order] += dS;
    }
    const double permanent = 4.4228e-8 * (-0.31460) * 0.30190;
    if (k.zeroTide) out.C[2][0] -= permanent;
    double m1, m2;
    poleWobble(in, m1, m2);
    out.C[2][1] += -1.333e-9 * (m1 + 0.0115 * m2);
    out.S[2][1] += -1.333e-9 * (m2 - 0.0115 * m1);
}

// HpopCorrections.h:302-305: Doodson arguments (tau, s, h, p, N', p_s) from
// the Delaunay arguments (l, l', F, D, Omega) and GMST.
inline void doodsonArguments(const double f[5], double gmst, double beta[6]) {
    const double s = f[2] + f[4], h = s - f[3];
    beta[0] = gmst + kPi - s;
    beta[1] = s;
    beta[2] = h;
    beta[3] = s - f[0];
    beta[4] = -f[4];
    beta[5] = h - f[1];
}

// HpopCorrections.h:306-317: the FES2004 wave sum added into out.
inline void oceanWaves(const double f[5], double gmst, const TideTables& tt, CS7& out) {
    double beta[6];
    doodsonArguments(f, gmst, beta);
    for (const OceanWave& w : tt.waves) {
        double theta = 0.0;
        for (int j = 0; j < 6; ++j) theta += w.doodson[j] * beta[j];
        const double co = std::cos(theta), si = std::sin(theta);
        for (int i = w.begin; i < w.end; ++i) {
            const OceanTerm& t = tt.terms[size_t(i)];
            out.C[t.n][t.m] += (t.cp + t.cm) * co + (t.sp + t.sm) * si;
            if (t.m) out.S[t.n][t.m] += (t.sp - t.sm) * co - (t.cp - t.cm) * si;
        }
    }
}

// HpopCorrections.h:298-322
inline void oceanCoefficients(const TideInput& in, const TideTables& tt, CS7& out) {
    ORACLE_FLOP_SCOPE(kEpoch);
    clearCS(out);
    double f[5], gmst;
    tideArguments(in, f, gmst);
    oceanWaves(f, gmst, tt, out);
    double m1, m2;
    poleWobble(in, m1, m2);
    out.C[2][1] += -2.1778e-10 * (m1 - 0.01724 * m2);
    out.S[2][1] += -1.7232e-10 * (m2 - 0.03365 * m1);
}

// HpopCorrections.h:339-367
inline void tideAcceleration(const TideInput& in, const TideConstants& k, const CS7& coeff, double out[3], int maxDegree) {
    double r[3];
    rot(in.E, in.r, r);
    const double d = std::sqrt(dot3(r, r));
    const double phi = std::asin(r[2] / d), lon = std::atan2(r[1], r[0]);
    double p[7][7], dp[7][7];
    legendre(maxDegree, phi, p, dp);
    double dUdr = 0.0, dUdlat = 0.0, dUdlon = 0.0;
    double cosM[7], sinM[7];
    for (int m = 0; m <= maxDegree; ++m) { cosM[m] = std::cos(m * lon); sinM[m] = std::sin(m * lon); }
    for (int n = 0; n <= maxDegree; ++n) {
        const double power = std::pow(k.radius / d, double(n));
        const double b1 = (-k.muEarth / (d * d)) * power * (n + 1);
        const double b2 = (k.muEarth / d) * power;
        double q1 = 0.0, q2 = 0.0, q3 = 0.0;
        for (int m = 0; m <= n; ++m) {
            const double value = coeff.C[n][m] * cosM[m] + coeff.S[n][m] * sinM[m];
            q1 += p[n][m] * value;
            q2 += dp[n][m] * value;
            q3 += m * p[n][m] * (coeff.S[n][m] * cosM[m] - coeff.C[n][m] * sinM[m]);
        }
        dUdr += q1 * b1;
        dUdlat += q2 * b2;
        dUdlon += q3 * b2;
    }
    const double xy2 = r[0] * r[0] + r[1] * r[1], xy = std::sqrt(xy2);
    const double q = (1.0 / d * dUdr - r[2] / (d * d * xy) * dUdlat);
    const double bf[3] = {q * r[0] - (1.0 / xy2 * dUdlon) * r[1], q * r[1] + (1.0 / xy2 * dUdlon) * r[0],
                          1.0 / d * dUdr * r[2] + xy / (d * d) * dUdlat};
    for (int i = 0; i < 3; ++i) out[i] = in.E[i] * bf[0] + in.E[3 + i] * bf[1] + in.E[6 + i] * bf[2];
}

// HpopCorrections.h:369-376
inline void relativity(const TideInput& in, const TideConstants& k, double out[3]) {
    const double r = std::sqrt(dot3(in.r, in.r));
    const double v = std::sqrt(dot3(in.v, in.v));
    const double rv = dot3(in.r, in.v);
    const double f = k.muEarth / (k.c * k.c * std::pow(r, 3.0));
    for (int i = 0; i < 3; ++i) out[i] = f * ((4.0 * k.muEarth / r - v * v) * in.r[i] + 4.0 * rv * in.v[i]);
}

}  // namespace oracle
