// Device functions of the epoch-only force-model inputs, shared by the
// epoch-record kernel (exact stage epochs) and the step kernel (per-object
// epochs on the uneven-epoch path).
//
// Reference behaviour reproduced (src/common/engine/paraHPOP/model/ unless noted):
//   EOP time model            environment/earthcorrections/EarthCorrections.h:46-105
//   tide arguments, GMST06    accelerations/hpop_compat/HpopCorrections.h:94-103,175-199
//   solid tide coefficients   HpopCorrections.h:148-169,201-264
//   ocean tide coefficients   HpopCorrections.h:298-322
//   NRLMSISE-00 epoch inputs  environment/atmosphere/Nrlmsise00.cu:91-241
// The analytic ephemeris and Earth orientation have no reference counterpart
// (the configurations use them in place of .brie/.brot2 files): Montenbruck &
// Gill sec. 3.3.2 Sun/Moon series, Standish (JPL) planetary elements, and the
// IERS 2010 CIO-based rotation with IAU 2006 precession polynomials.
#pragma once

#include <hip/hip_runtime.h>

#include "phx_common.h"
#include "phx_kernels.h"
#include "phx_nrlmsise.h"

namespace phx {
namespace epochdev {
namespace {


constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kDeg = kPi / 180.0;
constexpr double kArcsec = kPi / 648000.0;
constexpr double kAuKm = 149597870.7;
constexpr double kEarthRate = 7.292115146706979e-5;

struct D3 { double x, y, z; };

__device__ inline double posMod(double x, double p) { double a = fmod(x, p); return a < 0.0 ? a + p : a; }

__device__ double tdbMinusTt(double mjdTt) {
    const double t = (mjdTt - 51544.5) / 36525.0;
    return 0.001657 * sin(628.3076 * t + 6.2401) + 0.000022 * sin(575.3385 * t + 4.2970) +
           0.000014 * sin(1256.6152 * t + 6.1969) + 0.000005 * sin(606.9777 * t + 4.0212) +
           0.000005 * sin(52.9691 * t + 0.4444) + 0.000002 * sin(21.3299 * t + 5.5431) +
           0.000010 * sin(628.3076 * t + 4.2490);
}

struct Times { double mjdUtc, ut1Utc, ttUtc, xp, yp; bool valid; };

// Earth rotation angle offset (rad) that the leap seconds before `et` add to
// the analytic orientation: the EOP time model's UTC, and with it UT1 = UTC +
// (UT1-UTC), steps back by each leap second while the table's UT1-UTC stays
// continuous. leaps: [n] step epochs (ET, ascending), then [n] offsets.
__device__ inline double leapTheta(const double* leaps, int n, double et) {
    double th = 0.0;
    for (int k = 0; k < n; ++k) {
        if (et < leaps[k]) break;
        th = leaps[n + k];
    }
    return th;
}

// ET -> UTC/UT1/TT offsets and polar motion from the daily EOP table.
__device__ Times eopTimes(const EpochArgs& a, double et) {
    Times t;
    t.valid = false;
    t.mjdUtc = t.ut1Utc = t.ttUtc = t.xp = t.yp = nan("");
    if (a.eopDays < 2 || !isfinite(et)) return t;
    double ttEt = et;
    for (int j = 0; j < 2; ++j) ttEt = et - tdbMinusTt(51544.5 + ttEt / 86400.0);
    double utc = 51544.5 + (ttEt - 69.184) / 86400.0;
    for (int j = 0; j < 3; ++j) {
        double day = floor(utc - a.eopFirst);
        day = fmin(fmax(day, 0.0), double(a.eopDays - 1));
        utc = 51544.5 + (ttEt - 32.184 - a.eop[3 * a.eopDays + int(day)]) / 86400.0;
    }
    if (!(utc >= a.eopFirst) || !(utc < a.eopFirst + double(a.eopDays - 1))) return t;
    const int i = int(floor(utc - a.eopFirst));
    const double f = utc - floor(utc);
    const double* xp = a.eop;
    const double* yp = a.eop + a.eopDays;
    const double* du = a.eop + 2 * a.eopDays;
    const double* dat = a.eop + 3 * a.eopDays;
    t.mjdUtc = utc;
    t.ut1Utc = du[i] + f * (du[i + 1] - du[i]);
    t.ttUtc = 32.184 + dat[i];
    t.xp = xp[i] + f * (xp[i + 1] - xp[i]);
    t.yp = yp[i] + f * (yp[i + 1] - yp[i]);
    t.valid = true;
    return t;
}

__device__ double earthRotationAngle(double mjdUt1) {
    const double eraT = mjdUt1 + (2400000.5 - 2451545.0);
    const double f = fmod(mjdUt1, 1.0) + 0.5;
    return posMod(kTwoPi * (f + 0.7790572732640 + 0.00273781191135448 * eraT), kTwoPi);
}

__device__ void mat3mul(const double* A, const double* B, double* C) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    for (int i = 0; i < 9; ++i) C[i] = t[i];
}

__device__ void rotZ(double a, double* M) {
    const double c = cos(a), s = sin(a);
    M[0] = c; M[1] = s; M[2] = 0; M[3] = -s; M[4] = c; M[5] = 0; M[6] = 0; M[7] = 0; M[8] = 1;
}

// R = R1(-yp) R2(-xp) R3(s') * R3(ERA) * R3(-s) * [CIP matrix], omega = R^T (0, 0, w_E)
__device__ void orientation(const EpochArgs& a, double et, double* R, double* omega) {
    double mjdUt1, t, xp, yp;
    if (a.eopDays >= 2) {
        const Times tm = eopTimes(a, et);
        mjdUt1 = tm.mjdUtc + tm.ut1Utc / 86400.0;
        t = (tm.mjdUtc + tm.ttUtc / 86400.0 - 51544.5) / 36525.0;
        xp = tm.xp;
        yp = tm.yp;
    } else {
        double ttEt = et;
        for (int j = 0; j < 2; ++j) ttEt = et - tdbMinusTt(51544.5 + ttEt / 86400.0);
        mjdUt1 = 51544.5 + (ttEt - 69.184) / 86400.0;
        t = ttEt / (86400.0 * 36525.0);
        xp = yp = 0.0;
    }
    const double X = (-0.016617 + t * (2004.191898 + t * (-0.4297829 + t * (-0.19861834)))) * kArcsec;
    const double Y = (-0.006951 + t * (-0.025896 + t * (-22.4072747 + t * 0.00190059))) * kArcsec;
    const double s = -0.5 * X * Y + (0.000094 + 0.00380865 * t) * kArcsec;
    const double aa = 1.0 / (1.0 + sqrt(1.0 - X * X - Y * Y));
    double C[9] = {1 - aa * X * X, -aa * X * Y, -X, -aa * X * Y, 1 - aa * Y * Y, -Y, X, Y, 1 - aa * (X * X + Y * Y)};
    double M[9];
    // GCRS -> CIRS = R3(-s) R3(-E) R2(d) R3(E) (the transpose of IERS 2010 eq. 5.10's Q)
    rotZ(-s, M);
    mat3mul(M, C, C);
    rotZ(earthRotationAngle(mjdUt1), M);
    mat3mul(M, C, C);
    // TIRS -> ITRS = R1(-yp) R2(-xp) R3(s') (the transpose of eq. 5.3's W)
    const double sp = -47e-6 * t * kArcsec;
    const double cx = cos(xp * kArcsec), sx = sin(xp * kArcsec), cy = cos(yp * kArcsec), sy = sin(yp * kArcsec);
    double R2[9] = {cx, 0, sx, 0, 1, 0, -sx, 0, cx};
    double R1[9] = {1, 0, 0, 0, cy, -sy, 0, sy, cy};
    double W[9];
    rotZ(sp, M);
    mat3mul(R2, M, W);
    mat3mul(R1, W, W);
    mat3mul(W, C, R);
    omega[0] = R[6] * kEarthRate;
    omega[1] = R[7] * kEarthRate;
    omega[2] = R[8] * kEarthRate;
}

// ---- IERS2000 orientation from a .brot2 unit ----
// Lagrange value and derivative sweep over NP stencil nodes (lagrange/Interpolator.h:45-95):
// uniform grids in the local coordinate (nodes 0..NP-1), non-uniform ones over the keys.
template <int NP, int NCH, bool UNIFORM>
__device__ inline void lagrangeSweep(const double* __restrict__ vals, int n, int lo, double q,
                                     const double* __restrict__ keys, double* val, double* der) {
    for (int c = 0; c < NCH; ++c) val[c] = der[c] = 0.0;
    for (int j = 0; j < NP; ++j) {
        double prod = 1.0, dprod = 0.0, denom = 1.0;
        for (int m = 0; m < NP; ++m) {
            if (m == j) continue;
            double f, dfac;
            if (UNIFORM) {
                f = q - double(m);
                dfac = double(j - m);
            } else {
                const double xm = keys[lo + m];
                f = q - xm;
                dfac = keys[lo + j] - xm;
            }
            dprod = fma(dprod, f, prod);
            prod *= f;
            denom *= dfac;
        }
        const double invD = 1.0 / denom;
        const double wv = prod * invD, wd = dprod * invD;
        for (int c = 0; c < NCH; ++c) {
            const double y = vals[c * n + lo + j];
            val[c] = fma(wv, y, val[c]);
            der[c] = fma(wd, y, der[c]);
        }
    }
}

// ICRF -> ITRF R = PM ER PN and omega = vee(Rdot^T R) from the interpolated
// channels [X, Y, s, xp, yp, UT1-TT (days), dX, dY] and their per-day rates
// (orientations/detail/IERS2000Model.h:120-300,380-400), at the TT model time of
// the TDB epoch (:505-560).
__device__ void iersOrientation(const EpochArgs& a, double et, double* R, double* omega) {
    const double* d = a.iers;
    const int n1 = a.iersNut, n2 = a.iersErp;
    const double invSec = 1.0 / 86400.0;
    const double tdbMjd = et * invSec + 0.5;
    const double g = (kPi / 180.0) * (357.53 + 0.9856003 * (tdbMjd - 0.5));
    const double ttCorr = (0.001658 * sin(g) + 0.000014 * sin(2.0 * g)) * invSec;
    double frac = fmod(et, 86400.0) * invSec + 0.5;
    frac -= floor(frac);
    const double ttFrac = frac - ttCorr;
    const double tt = tdbMjd - ttCorr;
    const double* cst = d + 3 * n1 + 5 * n2;
    // nutation: uniform grid, six-node stencil centred on the bracketing interval
    const double x0 = cst[0], h = cst[1];
    long lo = long(floor((tt - x0) / h)) - 2;
    lo = lo < 0 ? 0 : (lo > n1 - 6 ? n1 - 6 : lo);
    const double u = (tt - (x0 + double(lo) * h)) / h;
    double nv[3], nd[3];
    lagrangeSweep<6, 3, true>(d, n1, int(lo), u, nullptr, nv, nd);
    const double invH = 1.0 / h;
    // ERP: node keys searched (largest key <= tt), four-node stencil
    const double* keys = cst + 4;
    int blo = 0, bhi = n2;
    while (blo < bhi) {
        const int mid = blo + ((bhi - blo) >> 1);
        if (keys[mid] <= tt) blo = mid + 1;
        else bhi = mid;
    }
    int st = blo - 1 - 1;
    st = st < 0 ? 0 : (st > n2 - 4 ? n2 - 4 : st);
    double ev[5], ed[5];
    lagrangeSweep<4, 5, false>(d + 3 * n1, n2, st, tt, keys, ev, ed);
    // prelude (:130-185)
    const double X = nv[0], Y = nv[1], s = nv[2];
    const double Xd = nd[0] * invH * invSec, Yd = nd[1] * invH * invSec, sd = nd[2] * invH * invSec;
    const double xp = ev[0], yp = ev[1], dut1 = ev[2], dX = ev[3], dY = ev[4];
    const double xpd = ed[0] * invSec, ypd = ed[1] * invSec, dut1d = ed[2] * invSec, dXd = ed[3] * invSec, dYd = ed[4] * invSec;
    const double x = X + dX, y = Y + dY, xdot = Xd + dXd, ydot = Yd + dYd;
    const double r2 = x * x + y * y, rr = sqrt(r2), fr = 1.0 / sqrt(1.0 - r2);
    const double e = r2 > 0.0 ? atan2(y, x) : 0.0;
    const double dd = atan(rr * fr);
    const double edot = r2 > 0.0 ? (ydot * x - xdot * y) / r2 : 0.0;
    const double ddot = (x * xdot + y * ydot) * fr / rr;
    const double mems = -e - s, memsdot = -edot - sd;
    const double ut1jd = tt + dut1 - 0.5;
    const double f = fmod(ttFrac + dut1 - 0.5, 1.0);
    const double twoPi = 2.0 * kPi;
    double theta = twoPi * (f + 0.7790572732640 + 0.00273781191135448 * ut1jd);
    theta = theta - twoPi * floor(theta / twoPi);
    const double thetadot = twoPi * ((invSec + dut1d) + 0.00273781191135448 * invSec);
    const double sp = -6.24e-15 * (tt - 0.5), spdot = -7.2205e-20;
    // matrices (:187-262)
    const double s1 = sin(e), c1 = cos(e), s2 = sin(dd), c2 = cos(dd), s3 = sin(mems), c3 = cos(mems);
    const double PN[9] = {c1 * c2 * c3 - s1 * s3, c1 * s3 + s1 * c2 * c3, -c3 * s2,
                          -s1 * c3 - c1 * c2 * s3, c1 * c3 - s1 * c2 * s3, s2 * s3,
                          s2 * c1, s2 * s1, c2};
    double PNd[9];
    {
        const double c1d = -s1 * edot, s1d = c1 * edot, c2d = -s2 * ddot, s2d = c2 * ddot, c3d = -s3 * memsdot, s3d = c3 * memsdot;
        const double c2s3 = c2 * s3, c2c3 = c2 * c3, c2s3d = c2d * s3 + c2 * s3d, c2c3d = c2d * c3 + c2 * c3d;
        PNd[0] = -s1d * s3 - s1 * s3d + c1d * c2c3 + c1 * c2c3d;
        PNd[1] = c1d * s3 + c1 * s3d + s1d * c2c3 + s1 * c2c3d;
        PNd[2] = -c3d * s2 - c3 * s2d;
        PNd[3] = -s1d * c3 - s1 * c3d - c1d * c2s3 - c1 * c2s3d;
        PNd[4] = c1d * c3 + c1 * c3d - s1d * c2s3 - s1 * c2s3d;
        PNd[5] = s2d * s3 + s2 * s3d;
        PNd[6] = s2d * c1 + s2 * c1d;
        PNd[7] = s2d * s1 + s2 * s1d;
        PNd[8] = c2d;
    }
    const double ct = cos(theta), stt = sin(theta);
    const double ER[9] = {ct, stt, 0.0, -stt, ct, 0.0, 0.0, 0.0, 1.0};
    const double ERd[9] = {-stt * thetadot, ct * thetadot, 0.0, -ct * thetadot, -stt * thetadot, 0.0, 0.0, 0.0, 0.0};
    const double PM[9] = {1.0, sp, xp, -sp, 1.0, -yp, -xp, yp, 1.0};
    const double PMd[9] = {0.0, spdot, xpd, -spdot, 0.0, -ypd, -xpd, ypd, 0.0};
    double ERPN[9], A[9], T1[9], T2[9], T3[9], Rd[9];
    mat3mul(ER, PN, ERPN);
    mat3mul(PM, ERPN, R);
    mat3mul(ERd, PN, A);
    mat3mul(PM, A, T1);
    mat3mul(PMd, ERPN, T2);
    mat3mul(ER, PNd, A);
    mat3mul(PM, A, T3);
    for (int i = 0; i < 9; ++i) Rd[i] = T1[i] + T2[i] + T3[i];
    double M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[3 * i + j] = Rd[i] * R[j] + Rd[3 + i] * R[3 + j] + Rd[6 + i] * R[6 + j];
    omega[0] = 0.5 * (M[7] - M[5]);
    omega[1] = 0.5 * (M[2] - M[6]);
    omega[2] = 0.5 * (M[3] - M[1]);
}

// ---- analytic ephemeris (geocentric, ICRF-aligned km) ----
__device__ D3 eclToEq(double x, double y, double z) {
    const double eps = 23.43929111 * kDeg;
    const double c = cos(eps), s = sin(eps);
    return D3{x, c * y - s * z, s * y + c * z};
}

__device__ double wrap360(double d) { d = fmod(d, 360.0); return d < 0 ? d + 360.0 : d; }

__device__ D3 sunPos(double T) {
    const double M = wrap360(357.5256 + 35999.049 * T) * kDeg;
    const double lon = 282.9400 * kDeg + M + (6892.0 * sin(M) + 72.0 * sin(2.0 * M)) * kArcsec;
    const double r = (149.619 - 2.499 * cos(M) - 0.021 * cos(2.0 * M)) * 1.0e6;
    return eclToEq(r * cos(lon), r * sin(lon), 0.0);
}

__device__ D3 moonPos(double T) {
    const double L0 = wrap360(218.31617 + 481267.88088 * T - 1.3972 * T) * kDeg;
    const double l = wrap360(134.96292 + 477198.86753 * T) * kDeg;
    const double lp = wrap360(357.52543 + 35999.04944 * T) * kDeg;
    const double F = wrap360(93.27283 + 483202.01873 * T) * kDeg;
    const double D = wrap360(297.85027 + 445267.11135 * T) * kDeg;
    const double dlon = 22640 * sin(l) + 769 * sin(2 * l) - 4586 * sin(l - 2 * D) + 2370 * sin(2 * D) - 668 * sin(lp) -
                        412 * sin(2 * F) - 212 * sin(2 * l - 2 * D) - 206 * sin(l + lp - 2 * D) + 192 * sin(l + 2 * D) -
                        165 * sin(lp - 2 * D) + 148 * sin(l - lp) - 125 * sin(D) - 110 * sin(l + lp) - 55 * sin(2 * F - 2 * D);
    const double lon = L0 + dlon * kArcsec;
    const double lat = (18520 * sin(F + lon - L0 + (412 * sin(2 * F) + 541 * sin(lp)) * kArcsec) - 526 * sin(F - 2 * D) +
                        44 * sin(l + F - 2 * D) - 31 * sin(-l + F - 2 * D) - 25 * sin(-2 * l + F) - 23 * sin(lp + F - 2 * D) +
                        21 * sin(-l + F) + 11 * sin(-lp + F - 2 * D)) * kArcsec;
    const double r = 385000 - 20905 * cos(l) - 3699 * cos(2 * D - l) - 2956 * cos(2 * D) - 570 * cos(2 * l) +
                     246 * cos(2 * l - 2 * D) - 205 * cos(lp - 2 * D) - 171 * cos(l + 2 * D) - 152 * cos(l + lp - 2 * D);
    return eclToEq(r * cos(lon) * cos(lat), r * sin(lon) * cos(lat), r * sin(lat));
}

// Standish (JPL) Table 1 elements for the planetary system barycenters 1..9 (3 = EMB unused).
static __constant__ double kPlanetEl[10][12] = {
    {0},
    {0.38709927, 0.20563593, 7.00497902, 252.25032350, 77.45779628, 48.33076593,
     0.00000037, 0.00001906, -0.00594749, 149472.67411175, 0.16047689, -0.12534081},
    {0.72333566, 0.00677672, 3.39467605, 181.97909950, 131.60246718, 76.67984255,
     0.00000390, -0.00004107, -0.00078890, 58517.81538729, 0.00268329, -0.27769418},
    {0},
    {1.52371034, 0.09339410, 1.84969142, -4.55343205, -23.94362959, 49.55953891,
     0.00001847, 0.00007882, -0.00813131, 19140.30268499, 0.44441088, -0.29257343},
    {5.20288700, 0.04838624, 1.30439695, 34.39644051, 14.72847983, 100.47390909,
     -0.00011607, -0.00013253, -0.00183714, 3034.74612775, 0.21252668, 0.20469106},
    {9.53667594, 0.05386179, 2.48599187, 49.95424423, 92.59887831, 113.66242448,
     -0.00125060, -0.00050991, 0.00193609, 1222.49362201, -0.41897216, -0.28867794},
    {19.18916464, 0.04725744, 0.77263783, 313.23810451, 170.95427630, 74.01692503,
     -0.00196176, -0.00004397, -0.00242939, 428.48202785, 0.40805281, 0.04240589},
    {30.06992276, 0.00859048, 1.77004347, -55.12002969, 44.96476227, 131.78422574,
     0.00026291, 0.00005105, 0.00035372, 218.45945325, -0.32241464, -0.00508664},
    {39.48211675, 0.24882730, 17.14001206, 238.92903833, 224.06891629, 110.30393684,
     -0.00031596, 0.00005170, 0.00004818, 145.20780515, -0.04062942, -0.01183482},
};

__device__ D3 planetHelio(int naif, double T) {
    double el[6];
    for (int k = 0; k < 6; ++k) el[k] = kPlanetEl[naif][k] + kPlanetEl[naif][6 + k] * T;
    const double a = el[0] * kAuKm, e = el[1], I = el[2] * kDeg;
    const double varpi = el[4] * kDeg, Om = el[5] * kDeg, w = varpi - Om;
    double M = fmod((el[3] - el[4]) * kDeg, kTwoPi);
    if (M > kPi) M -= kTwoPi;
    if (M < -kPi) M += kTwoPi;
    double E = M + e * sin(M);
    for (int it = 0; it < 6; ++it) E -= (E - e * sin(E) - M) / (1.0 - e * cos(E));
    const double xp = a * (cos(E) - e), yp = a * sqrt(1.0 - e * e) * sin(E);
    const double cw = cos(w), sw = sin(w), cO = cos(Om), sO = sin(Om), cI = cos(I), sI = sin(I);
    return eclToEq((cw * cO - sw * sO * cI) * xp + (-sw * cO - cw * sO * cI) * yp,
                   (cw * sO + sw * cO * cI) * xp + (-sw * sO + cw * cO * cI) * yp, (sw * sI) * xp + (cw * sI) * yp);
}

__device__ D3 bodyPos(int naif, double et) {
    const double T = et / (86400.0 * 36525.0);
    if (naif == 399) return D3{0, 0, 0};
    if (naif == 10) return sunPos(T);
    if (naif == 301) return moonPos(T);
    if (naif >= 1 && naif <= 9 && naif != 3) {
        const D3 p = planetHelio(naif, T), s = sunPos(T);
        return D3{p.x + s.x, p.y + s.y, p.z + s.z};
    }
    return D3{nan(""), nan(""), nan("")};
}

// ---- tides (SI), combined solid + ocean Stokes corrections ----
__device__ double gmst06(double mjdUt1, double mjdTt) {
    const double eraT = mjdUt1 + (2400000.5 - 2451545.0);
    const double f = fmod(mjdUt1, 1.0) + 0.5;
    const double era = posMod(kTwoPi * (f + 0.7790572732640 + 0.00273781191135448 * eraT), kTwoPi);
    const double t = ((2400000.5 - 2451545.0) + mjdTt) / 36525.0;
    return posMod(era + (0.014506 + (4612.156534 + (1.3915817 + (-0.00000044 + (-0.000029956 - 0.0000000368 * t) * t) * t) * t) * t) * kArcsec, kTwoPi);
}

__device__ void tideArgs(const Times& tm, double args[5], double& thetaG) {
    const double mjdTt = tm.mjdUtc + tm.ttUtc / 86400.0;
    const double t = (mjdTt - 51544.5) / 36525.0, t2 = t * t, t3 = t2 * t, t4 = t3 * t;
    args[0] = posMod(485868.249036 + 1717915923.2178 * t + 31.8792 * t2 + 0.051635 * t3 - 0.00024470 * t4, 1296000.0) * kArcsec;
    args[1] = posMod(1287104.79305 + 129596581.0481 * t - 0.5532 * t2 + 0.000136 * t3 - 0.00001149 * t4, 1296000.0) * kArcsec;
    args[2] = posMod(335779.526232 + 1739527262.8478 * t - 12.7512 * t2 - 0.001037 * t3 + 0.00000417 * t4, 1296000.0) * kArcsec;
    args[3] = posMod(1072260.70369 + 1602961601.2090 * t - 6.3706 * t2 + 0.006593 * t3 - 0.00003169 * t4, 1296000.0) * kArcsec;
    args[4] = posMod(450160.398036 - 6962890.5431 * t + 7.4722 * t2 + 0.007702 * t3 - 0.00005939 * t4, 1296000.0) * kArcsec;
    thetaG = gmst06(tm.mjdUtc + tm.ut1Utc / 86400.0, mjdTt);
}

__device__ void poleWobble(const Times& tm, double& m1, double& m2) {
    const double t = (tm.mjdUtc + tm.ttUtc / 86400.0 - 51544.5) / 365.25;
    double x, y;
    if (t < 10.0) {
        x = 55.974 + t * (1.8243 + t * (0.18413 + t * 0.007024));
        y = 346.346 + t * (1.7896 + t * (-0.10729 - t * 0.000908));
    } else {
        x = 23.513 + 7.6141 * t;
        y = 358.891 - 0.6287 * t;
    }
    m1 = tm.xp - 1e-3 * x;
    m2 = -(tm.yp - 1e-3 * y);
}

// Fully normalized P_nm (n <= 3) at geocentric latitude with sin/cos given.
__device__ void legendre3(double s, double c, double p[4][4]) {
    for (int n = 0; n < 4; ++n) for (int m = 0; m < 4; ++m) p[n][m] = 0.0;
    p[0][0] = 1.0;
    p[1][1] = sqrt(3.0) * c;
    for (int n = 2; n <= 3; ++n) p[n][n] = sqrt((2.0 * n + 1) / (2.0 * n)) * c * p[n - 1][n - 1];
    for (int n = 1; n <= 3; ++n) p[n][n - 1] = sqrt(2.0 * n + 1) * s * p[n - 1][n - 1];
    for (int m = 0; m <= 3; ++m)
        for (int n = m + 2; n <= 3; ++n)
            p[n][m] = sqrt((2.0 * n + 1) / ((n - m) * double(n + m))) *
                      (sqrt(2.0 * n - 1) * s * p[n - 1][m] - sqrt(((n + m - 1) * double(n - m - 1)) / (2.0 * n - 3)) * p[n - 2][m]);
}

__device__ void solidTide(const EpochArgs& a, const Times& tm, const double* E, D3 sunKm, D3 moonKm, double C[7][7], double S[7][7]) {
    const double muEarth = 398600.4415e9, radius = 6378.1363e3, muSun = 132712440041.279419e9,
                 muMoon = 398600.4415e9 / 81.3005682214972154;
    double moon[3], sun[3];
    const double mi[3] = {1000.0 * moonKm.x, 1000.0 * moonKm.y, 1000.0 * moonKm.z};
    const double si[3] = {1000.0 * sunKm.x, 1000.0 * sunKm.y, 1000.0 * sunKm.z};
    for (int i = 0; i < 3; ++i) {
        moon[i] = E[3 * i] * mi[0] + E[3 * i + 1] * mi[1] + E[3 * i + 2] * mi[2];
        sun[i] = E[3 * i] * si[0] + E[3 * i + 1] * si[1] + E[3 * i + 2] * si[2];
    }
    const double mLon = atan2(moon[1], moon[0]), sLon = atan2(sun[1], sun[0]);
    const double mr = sqrt(moon[0] * moon[0] + moon[1] * moon[1] + moon[2] * moon[2]);
    const double sr = sqrt(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]);
    const double mRatio = radius / mr, sRatio = radius / sr;
    double mP[4][4], sP[4][4];
    const double mxy = hypot(moon[0], moon[1]), sxy = hypot(sun[0], sun[1]);
    legendre3(moon[2] / mr, mxy / mr, mP);
    legendre3(sun[2] / sr, sxy / sr, sP);
    const double mF = muMoon / muEarth * mRatio * mRatio * mRatio;
    const double sF = muSun / muEarth * sRatio * sRatio * sRatio;
    double qC[3], qS[3];
    for (int m = 0; m <= 2; ++m) {
        qC[m] = mF * mP[2][m] * cos(m * mLon) + sF * sP[2][m] * cos(m * sLon);
        qS[m] = mF * mP[2][m] * sin(m * mLon) + sF * sP[2][m] * sin(m * sLon);
    }
    C[2][0] += 0.30190 / 5.0 * qC[0];
    C[2][1] += 0.29830 / 5.0 * qC[1] - 0.00144 / 5.0 * qS[1];
    S[2][1] += 0.00144 / 5.0 * qC[1] + 0.29830 / 5.0 * qS[1];
    C[2][2] += 0.30102 / 5.0 * qC[2] - 0.00130 / 5.0 * qS[2];
    S[2][2] += 0.00130 / 5.0 * qC[2] + 0.30102 / 5.0 * qS[2];
    C[4][0] += -0.00089 / 5.0 * qC[0];
    C[4][1] += -0.00080 / 5.0 * qC[1];
    S[4][1] += -0.00080 / 5.0 * qS[1];
    C[4][2] += -0.00057 / 5.0 * qC[2];
    S[4][2] += -0.00057 / 5.0 * qS[2];
    const double m3 = muMoon / muEarth * mRatio * mRatio * mRatio * mRatio;
    const double s3 = muSun / muEarth * sRatio * sRatio * sRatio * sRatio;
    for (int m = 0; m <= 3; ++m) {
        const double f = (m == 3 ? 0.094 : 0.093) / 7.0;
        C[3][m] += f * (m3 * mP[3][m] * cos(m * mLon) + s3 * sP[3][m] * cos(m * sLon));
        if (m) S[3][m] += f * (m3 * mP[3][m] * sin(m * mLon) + s3 * sP[3][m] * sin(m * sLon));
    }
    double args[5], thetaG;
    tideArgs(tm, args, thetaG);
    int idx = 0;
    for (int order = 0; order <= 2; ++order) {
        double dC = 0.0, dS = 0.0;
        const int count = a.freqCount[order];
        for (int i = 0; i < count; ++i, ++idx) {
            const double* row = a.freq + 7 * idx;  // 5 multipliers, in-phase, out-of-phase
            double arg = 0.0;
            for (int j = 0; j < 5; ++j) arg += row[j] * args[j];
            const double ph = order * (thetaG + kPi) - arg;
            const double c = cos(ph), s = sin(ph);
            if (order == 0) dC += 1e-12 * (row[5] * c - row[6] * s);
            else if (order == 1) {
                dC += 1e-12 * (row[5] * s + row[6] * c);
                dS += 1e-12 * (row[5] * c - row[6] * s);
            } else {
                dC += 1e-12 * row[5] * c;
                dS -= 1e-12 * row[5] * s;
            }
        }
        C[2][order] += dC;
        S[2][order] += dS;
    }
    if (a.zeroTide) C[2][0] -= 4.4228e-8 * (-0.31460) * 0.30190;
    double m1, m2;
    poleWobble(tm, m1, m2);
    C[2][1] += -1.333e-9 * (m1 + 0.0115 * m2);
    S[2][1] += -1.333e-9 * (m2 - 0.0115 * m1);
}

__device__ void oceanTide(const EpochArgs& a, const Times& tm, double C[7][7], double S[7][7]) {
    double f[5], gmst;
    tideArgs(tm, f, gmst);
    const double s = f[2] + f[4], h = s - f[3];
    const double beta[6] = {gmst + kPi - s, s, h, s - f[0], -f[4], h - f[1]};
    for (int w = 0; w < a.waveCount; ++w) {
        const double* wv = a.waves + 8 * w;  // 6 Doodson multipliers, first term, end term
        double theta = 0.0;
        for (int j = 0; j < 6; ++j) theta += wv[j] * beta[j];
        const double co = cos(theta), si = sin(theta);
        for (int i = int(wv[6]); i < int(wv[7]); ++i) {
            const double* t = a.terms + 6 * i;   // n, m, Cp, Sp, Cm, Sm
            const int n = int(t[0]), m = int(t[1]);
            C[n][m] += (t[2] + t[4]) * co + (t[3] + t[5]) * si;
            if (m) S[n][m] += (t[3] - t[5]) * co - (t[2] - t[4]) * si;
        }
    }
    double m1, m2;
    poleWobble(tm, m1, m2);
    C[2][1] += -2.1778e-10 * (m1 - 0.01724 * m2);
    S[2][1] += -1.7232e-10 * (m2 - 0.03365 * m1);
}

// ---- NRLMSISE-00 epoch inputs (Nrlmsise00.cu:91-241) ----
__device__ double taiMinusUtcNrl(double mjd) {
    const double dates[27] = {41499, 41683, 42048, 42413, 42778, 43144, 43509, 43874, 44239, 44786, 45151, 45516, 46247, 47161,
                              47892, 48257, 48804, 49169, 49534, 50083, 50630, 51179, 53736, 54832, 56109, 57204, 57754};
    double dat = 10.0;
    for (int i = 0; i < 27; ++i)
        if (mjd >= dates[i]) dat = 11.0 + i;
    return dat;
}

__device__ double etToMjdUtcNrl(double et) {
    const double days = et / 86400.0;
    const double g = 6.240075674 + 0.01720197034 * days;
    const double dtt = 0.001657 * sin(g) + 0.000022 * sin(2.0 * g);
    const double mjdTt = 51544.5 + (et - dtt) / 86400.0;
    const double first = mjdTt - (32.184 + taiMinusUtcNrl(mjdTt)) / 86400.0;
    return mjdTt - (32.184 + taiMinusUtcNrl(first)) / 86400.0;
}

__device__ double wrapHours(double h) { h = fmod(h, 24.0); return h < 0.0 ? h + 24.0 : h; }

__device__ double gastHours(double mjdUtc) {
    const double mjd0 = floor(mjdUtc);
    const double ut1 = 86400.0 * (mjdUtc - mjd0);
    const double t0 = (mjd0 - 51544.5) / 36525.0, t = (mjdUtc - 51544.5) / 36525.0;
    const double gs = 24110.54841 + 8640184.812866 * t0 + 1.002737909350795 * ut1 + (0.093104 - 6.2e-6 * t) * t * t;
    const double gr = kTwoPi * (gs / 86400.0 - floor(gs / 86400.0));
    const double mjdTt = mjdUtc + (32.184 + taiMinusUtcNrl(mjdUtc)) / 86400.0;
    const double tt = (mjdTt - 51544.5) / 36525.0;
    const double om = (125.04452 - 1934.136261 * tt) * kDeg;
    const double sl = (280.4665 + 36000.7698 * tt) * kDeg;
    const double ml = (218.3165 + 481267.8813 * tt) * kDeg;
    const double dpsi = -17.20 * sin(om) - 1.32 * sin(2.0 * sl) - 0.23 * sin(2.0 * ml) + 0.21 * sin(2.0 * om);
    const double eps = (23.43929111 - (46.8150 + (0.00059 - 0.001813 * tt) * tt) * tt / 3600.0) * kDeg;
    return wrapHours((gr + dpsi / 3600.0 * kDeg * cos(eps)) * 24.0 / kTwoPi);
}

// Calendar day of year and UTC seconds of day of an MJD UTC
// (Nrlmsise00.cu:113-140).
__device__ inline void doySecOf(double mjd, int& doyOut, double& secOut) {
    const long A = long(floor(mjd + 2400001.0));
    long c;
    if (A < 2299161) c = A + 1524;
    else {
        const long b = long((A - 1867216.25) / 36524.25);
        c = A + b - b / 4 + 1525;
    }
    const long d = long((c - 122.1) / 365.25);
    const long e = 365 * d + d / 4;
    const long f = long((c - e) / 30.6001);
    const int day = int(c - e - long(30.6001 * f));
    const int month = int(f - 1 - 12 * (f / 14));
    const int year = int(d - 4715 - (7 + month) / 10);
    double sod = (mjd - floor(mjd)) * 86400.0;
    if (sod < 0.0) sod += 86400.0;
    const int hour = int(sod / 3600.0);
    const int minute = int((sod - hour * 3600.0) / 60.0);
    const double second = sod - hour * 3600.0 - minute * 60.0;
    const int dm[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    const bool leap = (year % 4 == 0 && year % 100 != 0) || year % 400 == 0;
    int doy = day;
#pragma unroll
    for (int m = 1; m < 12; ++m)
        if (m < month) doy += (m == 2 && leap) ? 29 : dm[m - 1];
    doyOut = doy;
    secOut = hour * 3600.0 + minute * 60.0 + second;
}

// Weather-table record and 3-hour slot of an MJD UTC; false outside the
// table's usable range (three days of history are needed).
__device__ inline bool weatherSlotOf(const EpochArgs& a, double mjd, int& record, int& slot) {
    record = int(floor(mjd - a.wStart));
    slot = int(floor((mjd - floor(mjd)) * 8.0));
    slot = slot < 0 ? 0 : (slot > 7 ? 7 : slot);
    return a.wDays > 0 && record >= 3 && record < a.wDays;
}

// Flux and Ap inputs of one weather slot (adjusted flux; Nrlmsise00.cu:215-241):
// out[0] F10.7 (previous day), out[1] F10.7A, out[2..8] the Ap array.
__device__ inline void weatherInputs(const EpochArgs& a, int record, int slot, double* out) {
    const double* W = a.weather;
    const int nd = a.wDays;
    auto at = [&](int comp, int r) { return W[comp * nd + r]; };
    auto lag = [&](int l) { const int h = 24 + slot - l; return at(5 + h % 8, record - 3 + h / 8); };
    out[0] = at(2, record - 1);
    out[1] = at(3, record);
    out[2] = at(4, record);
    out[3] = at(5 + slot, record);
    out[4] = lag(1);
    out[5] = lag(2);
    out[6] = lag(3);
    double s5 = 0.0, s6 = 0.0;
    for (int k = 4; k <= 11; ++k) s5 += lag(k);
    for (int k = 12; k <= 19; ++k) s6 += lag(k);
    out[7] = s5 / 8.0;
    out[8] = s6 / 8.0;
}

// Writes the NRLMSISE-00 epoch inputs at dst[REC_NRL_* - REC_NRL] and, when
// globe is given, the epoch-only globe7 terms (nrlmsise::GE_COUNT values)
// followed by the lower-thermosphere glob7s terms (nrlmsise::GL_COUNT values).
__device__ void nrlEpoch(const EpochArgs& a, double et, double* dst, double* globe) {
    double* rec = dst - REC_NRL;
    const double mjd = etToMjdUtcNrl(et);
    int doy;
    double sec;
    doySecOf(mjd, doy, sec);
    rec[REC_NRL_DOY] = double(doy);
    rec[REC_NRL_SEC] = sec;
    rec[REC_NRL_GAST] = gastHours(mjd);
    int record, slot;
    if (!weatherSlotOf(a, mjd, record, slot)) {
        for (int k = 0; k < 10; ++k) rec[REC_NRL_F107 + k] = nan("");
        rec[REC_NRL_VALID] = 0.0;
        return;
    }
    weatherInputs(a, record, slot, rec + REC_NRL_F107);
    rec[REC_NRL_VALID] = 1.0;
    if (globe && a.nrlCoeff) {
        nrlmsise::NrlTables T = nrlmsise::nrlTables(a.nrlCoeff);
        nrlmsise::globeEpoch(T, doy, rec[REC_NRL_SEC], rec[REC_NRL_F107], rec[REC_NRL_F107A], rec + REC_NRL_AP, globe);
        nrlmsise::lowerEpoch(T, doy, globe + nrlmsise::GE_COUNT);
    }
}

// One row of the weather-slot table (3-hour UTC slot `slot` of weather record
// `record`): the slot's flux/Ap inputs and the globe7 terms that are constant
// over the slot (nrlmsise::globeSlot).
__device__ void nrlSlotRow(const EpochArgs& a, int record, int slot, double* row) {
    using namespace nrlmsise;
    for (int k = 0; k < SLOT_SIZE; ++k) row[k] = 0.0;
    if (!(a.wDays > 0 && record >= 3 && record < a.wDays) || !a.nrlCoeff) return;
    const double mjd = a.wStart + double(record) + (double(slot) + 0.5) / 8.0;
    int doy;
    double sec;
    doySecOf(mjd, doy, sec);
    weatherInputs(a, record, slot, row + SLOT_F107);
    NrlTables T = nrlTables(a.nrlCoeff);
    globeSlot(T, doy, row[SLOT_F107], row[SLOT_F107A], row + SLOT_AP, row);
    row[SLOT_DOY] = double(doy);
    row[SLOT_VALID] = 1.0;
}

}  // namespace
}  // namespace epochdev
}  // namespace phx
