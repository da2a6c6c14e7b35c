// ORACLE — test infrastructure only. CPU restatement of the reference's
// NRLMSISE-00 density path:
//   model/environment/atmosphere/Nrlmsise00.cu:91-327   input preparation and
//       evaluateNrlmsise00HostDensity (ET->UTC, geodetic, weather, local time)
//   model/environment/atmosphere/detail/Nrlmsise00Model.inl   the model
//       (tselec, glatf, ccor, ccor2, scalh, dnet, splini, splint, spline,
//        zeta, densm, densu, g0, sumex, sg0, globe7, glob7s, gtd7, gtd7d, gts7)
// The coefficient blocks are read from nrlmsise00_coefficients.txt.
// Pinned by the 15 published test-driver outputs and the ap-array case
// (tests/golden/nrlmsise00_testdriver.json); the input preparation (ET -> UTC,
// day of year, seconds, GAST hours, weather record and 3-hour slots) by the
// published input definitions (tests/test_oracle_pins_env.py).
#pragma once

#include "flop_scope.h"

#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle_physics.h"

namespace oracle {

struct NrlmsiseTables {
    double pt[150], pd[9][150], ps[150], pdl[2][25], ptm[10], pdm[8][10], ptl[4][100], pma[10][100], sam[100], pavgm[10];
};

inline void loadNrlmsiseTables(const std::string& path, NrlmsiseTables& t) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("missing NRLMSISE-00 coefficient file: " + path);
    std::map<std::string, std::pair<double*, size_t>> dst = {
        {"pt", {t.pt, 150}}, {"pd", {&t.pd[0][0], 9 * 150}}, {"ps", {t.ps, 150}}, {"pdl", {&t.pdl[0][0], 50}},
        {"ptm", {t.ptm, 10}}, {"pdm", {&t.pdm[0][0], 80}}, {"ptl", {&t.ptl[0][0], 400}},
        {"pma", {&t.pma[0][0], 1000}}, {"sam", {t.sam, 100}}, {"pavgm", {t.pavgm, 10}}};
    std::string line, cur;
    size_t filled = 0;
    size_t blocks = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (std::isalpha((unsigned char)line[0])) {
            std::istringstream s(line);
            s >> cur;
            if (!dst.count(cur)) throw std::runtime_error("unknown NRLMSISE-00 block " + cur);
            filled = 0;
            ++blocks;
            continue;
        }
        std::istringstream s(line);
        double v;
        while (s >> v) {
            auto& d = dst.at(cur);
            if (filled >= d.second) throw std::runtime_error("too many values in block " + cur);
            d.first[filled++] = v;
        }
    }
    if (blocks != dst.size()) throw std::runtime_error("incomplete NRLMSISE-00 coefficient file");
}

struct NrlWeather {
    double startMjdUtc = 0;
    int days = 0;
    std::vector<double> data;  // 13 x days, SoA
    double at(int comp, int rec) const { return data[size_t(comp) * days + rec]; }
};

namespace nrl {

struct Input {
    int doy;
    double sec, alt, g_lat, g_long, lst, f107A, f107, ap;
    double apa[7];
};

struct Flags {
    int switches[24];
    int sw[24];
    int swc[24];
};

struct Output { double d[9]; double t[2]; };

// Per-evaluation state (the model's file-scope variables in the C package).
struct State {
    double gsurf = 0, re = 0;
    double dd = 0;
    double dm04 = 0, dm16 = 0, dm28 = 0, dm32 = 0, dm40 = 0, dm01 = 0, dm14 = 0;
    double meso_tn1[5] = {}, meso_tn2[4] = {}, meso_tn3[5] = {};
    double meso_tgn1[2] = {}, meso_tgn2[2] = {}, meso_tgn3[2] = {};
    double dfa = 0;
    double plg[4][9] = {};
    double ctloc = 0, stloc = 0, c2tloc = 0, s2tloc = 0, s3tloc = 0, c3tloc = 0;
    double apdf = 0, apt[4] = {};
    bool g0Valid = false;
    double g0P24 = 0, g0P25 = 0, g0Cache[6] = {};
};

// Nrlmsise00Model.inl:80-100
inline void tselec(Flags& f) {
    for (int i = 0; i < 24; ++i) {
        if (i != 9) {
            f.sw[i] = f.switches[i] == 1 ? 1 : 0;
            f.swc[i] = f.switches[i] > 0 ? 1 : 0;
        } else {
            f.sw[i] = f.switches[i];
            f.swc[i] = f.switches[i];
        }
    }
}

// :108-114
inline void glatf(double lat, double& gv, double& reff) {
    const double dgtr = 1.74533E-2;
    const double c2 = std::cos(2.0 * dgtr * lat);
    gv = 980.616 * (1.0 - 0.0026373 * c2);
    reff = 2.0 * gv / (3.085462E-6 + 2.27E-9 * c2) * 1.0E-5;
}

// :120-137
inline double ccor(double alt, double r, double h1, double zh) {
    double e = (alt - zh) / h1;
    if (e > 70) return std::exp(0.0);
    if (e < -70) return std::exp(r);
    const double ex = std::exp(e);
    e = r / (1.0 + ex);
    return std::exp(e);
}

// :145-164
inline double ccor2(double alt, double r, double h1, double zh, double h2) {
    const double e1 = (alt - zh) / h1, e2 = (alt - zh) / h2;
    if (e1 > 70 || e2 > 70) return std::exp(0.0);
    if (e1 < -70 && e2 < -70) return std::exp(r);
    const double ex1 = std::exp(e1), ex2 = std::exp(e2);
    return std::exp(r / (1.0 + 0.5 * (ex1 + ex2)));
}

// :172-178
inline double scalh(const State& st, double alt, double xm, double temp) {
    const double rgas = 831.4;
    double g = st.gsurf / std::pow((1.0 + alt / st.re), 2.0);
    g = rgas * temp / (g * xm);
    return g;
}

// :186-215
inline double dnet(double dd, double dm, double zhm, double xmm, double xm) {
    double a = zhm / (xmm - xm);
    if (!((dm > 0) && (dd > 0))) {
        if (dd == 0 && dm == 0) dd = 1;
        if (dm == 0) return dd;
        if (dd == 0) return dm;
    }
    const double ylog = a * std::log(dm / dd);
    if (ylog < -10) return dd;
    if (ylog > 10) return dm;
    a = dd * std::pow((1.0 + std::exp(ylog)), (1.0 / a));
    return a;
}

// :223-252
inline void splini(const double* xa, const double* ya, const double* y2a, int n, double x, double* y) {
    double yi = 0;
    int klo = 0, khi = 1;
    while ((x > xa[klo]) && (khi < n)) {
        double xx = x;
        if (khi < (n - 1)) xx = x < xa[khi] ? x : xa[khi];
        const double h = xa[khi] - xa[klo];
        const double a = (xa[khi] - xx) / h;
        const double b = (xx - xa[klo]) / h;
        const double a2 = a * a, b2 = b * b;
        yi += ((1.0 - a2) * ya[klo] / 2.0 + b2 * ya[khi] / 2.0 +
               ((-(1.0 + a2 * a2) / 4.0 + a2 / 2.0) * y2a[klo] + (b2 * b2 / 4.0 - b2 / 2.0) * y2a[khi]) * h * h / 6.0) * h;
        klo++;
        khi++;
    }
    *y = yi;
}

// :260-290
inline void splint(const double* xa, const double* ya, const double* y2a, int n, double x, double* y) {
    int klo = 0, khi = n - 1;
    while ((khi - klo) > 1) {
        const int k = (khi + klo) / 2;
        if (xa[k] > x) khi = k;
        else klo = k;
    }
    const double h = xa[khi] - xa[klo];
    const double a = (xa[khi] - x) / h;
    const double b = (x - xa[klo]) / h;
    *y = a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * h * h / 6.0;
}

// :298-338
inline void spline(const double* x, const double* y, int n, double yp1, double ypn, double* y2) {
    double u[5];
    if (yp1 > 0.99E30) {
        y2[0] = 0;
        u[0] = 0;
    } else {
        y2[0] = -0.5;
        u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
    for (int i = 1; i < (n - 1); ++i) {
        const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]) -
                sig * u[i - 1]) / p;
    }
    double qn, un;
    if (ypn > 0.99E30) {
        qn = 0;
        un = 0;
    } else {
        qn = 0.5;
        un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    }
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

// :346-348
inline double zeta(const State& st, double zz, double zl) { return ((zz - zl) * (st.re + zl) / (st.re + zz)); }

// :350-468
inline double densm(State& st, double alt, double d0, double xm, double* tz, int mn3, const double* zn3, const double* tn3,
                    const double* tgn3, int mn2, const double* zn2, const double* tn2, const double* tgn2) {
    double xs[5], ys[5], y2out[5];
    const double rgas = 831.4;
    double densm_tmp = d0;
    if (alt > zn2[0]) return xm == 0.0 ? *tz : d0;
    double z = alt > zn2[mn2 - 1] ? alt : zn2[mn2 - 1];
    int mn = mn2;
    double z1 = zn2[0], z2 = zn2[mn - 1], t1 = tn2[0], t2 = tn2[mn - 1];
    double zg = zeta(st, z, z1), zgdif = zeta(st, z2, z1);
    for (int k = 0; k < mn; ++k) {
        xs[k] = zeta(st, zn2[k], z1) / zgdif;
        ys[k] = 1.0 / tn2[k];
    }
    double yd1 = -tgn2[0] / (t1 * t1) * zgdif;
    double yd2 = -tgn2[1] / (t2 * t2) * zgdif * (std::pow(((st.re + z2) / (st.re + z1)), 2.0));
    spline(xs, ys, mn, yd1, yd2, y2out);
    double x = zg / zgdif, y;
    splint(xs, ys, y2out, mn, x, &y);
    *tz = 1.0 / y;
    if (xm != 0.0) {
        const double glb = st.gsurf / (std::pow((1.0 + z1 / st.re), 2.0));
        const double gamm = xm * glb * zgdif / rgas;
        double yi;
        splini(xs, ys, y2out, mn, x, &yi);
        double expl = gamm * yi;
        if (expl > 50.0) expl = 50.0;
        densm_tmp = densm_tmp * (t1 / *tz) * std::exp(-expl);
    }
    if (alt > zn3[0]) return xm == 0.0 ? *tz : densm_tmp;
    z = alt;
    mn = mn3;
    z1 = zn3[0];
    z2 = zn3[mn - 1];
    t1 = tn3[0];
    t2 = tn3[mn - 1];
    zg = zeta(st, z, z1);
    zgdif = zeta(st, z2, z1);
    for (int k = 0; k < mn; ++k) {
        xs[k] = zeta(st, zn3[k], z1) / zgdif;
        ys[k] = 1.0 / tn3[k];
    }
    yd1 = -tgn3[0] / (t1 * t1) * zgdif;
    yd2 = -tgn3[1] / (t2 * t2) * zgdif * (std::pow(((st.re + z2) / (st.re + z1)), 2.0));
    spline(xs, ys, mn, yd1, yd2, y2out);
    x = zg / zgdif;
    splint(xs, ys, y2out, mn, x, &y);
    *tz = 1.0 / y;
    if (xm != 0.0) {
        const double glb = st.gsurf / (std::pow((1.0 + z1 / st.re), 2.0));
        const double gamm = xm * glb * zgdif / rgas;
        double yi;
        splini(xs, ys, y2out, mn, x, &yi);
        double expl = gamm * yi;
        if (expl > 50.0) expl = 50.0;
        densm_tmp = densm_tmp * (t1 / *tz) * std::exp(-expl);
    }
    return xm == 0.0 ? *tz : densm_tmp;
}

// :476-576
inline double densu(State& st, double alt, double dlb, double tinf, double tlb, double xm, double alpha, double* tz,
                    double zlb, double s2, int mn1, const double* zn1, double* tn1, double* tgn1) {
    double x = 0, y;
    const double rgas = 831.4;
    double densu_temp = 1.0;
    double z1 = 0, t1 = 0, zgdif = 0;
    int mn = 0;
    double xs[5], ys[5], y2out[5];
    const double za = zn1[0];
    double z = alt > za ? alt : za;
    const double zg2 = zeta(st, z, zlb);
    const double tt = tinf - (tinf - tlb) * std::exp(-s2 * zg2);
    const double ta = tt;
    *tz = tt;
    densu_temp = *tz;
    if (alt < za) {
        const double dta = (tinf - ta) * s2 * std::pow(((st.re + zlb) / (st.re + za)), 2.0);
        tgn1[0] = dta;
        tn1[0] = ta;
        z = alt > zn1[mn1 - 1] ? alt : zn1[mn1 - 1];
        mn = mn1;
        z1 = zn1[0];
        const double z2 = zn1[mn - 1];
        t1 = tn1[0];
        const double t2 = tn1[mn - 1];
        const double zg = zeta(st, z, z1);
        zgdif = zeta(st, z2, z1);
        for (int k = 0; k < mn; ++k) {
            xs[k] = zeta(st, zn1[k], z1) / zgdif;
            ys[k] = 1.0 / tn1[k];
        }
        const double yd1 = -tgn1[0] / (t1 * t1) * zgdif;
        const double yd2 = -tgn1[1] / (t2 * t2) * zgdif * std::pow(((st.re + z2) / (st.re + z1)), 2.0);
        spline(xs, ys, mn, yd1, yd2, y2out);
        x = zg / zgdif;
        splint(xs, ys, y2out, mn, x, &y);
        *tz = 1.0 / y;
        densu_temp = *tz;
    }
    if (xm == 0) return densu_temp;
    double glb = st.gsurf / std::pow((1.0 + zlb / st.re), 2.0);
    const double gamma = xm * glb / (s2 * rgas * tinf);
    double expl = std::exp(-s2 * gamma * zg2);
    if (expl > 50.0) expl = 50.0;
    if (tt <= 0) expl = 50.0;
    const double densa = dlb * std::pow((tlb / tt), ((1.0 + alpha + gamma))) * expl;
    densu_temp = densa;
    if (alt >= za) return densu_temp;
    glb = st.gsurf / std::pow((1.0 + z1 / st.re), 2.0);
    const double gamm = xm * glb * zgdif / rgas;
    double yi;
    splini(xs, ys, y2out, mn, x, &yi);
    expl = gamm * yi;
    if (expl > 50.0) expl = 50.0;
    if (*tz <= 0) expl = 50.0;
    densu_temp = densu_temp * std::pow((t1 / *tz), (1.0 + alpha)) * std::exp(-expl);
    return densu_temp;
}

// :586-593 (p[24] lower bound applied to a local, as in the reference)
inline double g0(double a, const double* p) {
    const double p24 = (p[24] < 1.0E-4) ? 1.0E-4 : std::sqrt(p[24] * p[24]);
    return (a - 4.0 + (p[25] - 1.0) * (a - 4.0 + (std::exp(-p24 * (a - 4.0)) - 1.0) / p24));
}

// :596-598
inline double sumex(double ex) { return (1.0 + (1.0 - std::pow(ex, 19.0)) / (1.0 - ex) * std::pow(ex, 0.5)); }

// :601-634
inline double sg0(State& st, double ex, const double* p, const double* ap) {
    double g[6];
    if (st.g0Valid && p[24] == st.g0P24 && p[25] == st.g0P25) {
        for (int k = 0; k < 6; ++k) g[k] = st.g0Cache[k];
    } else {
        for (int k = 0; k < 6; ++k) g[k] = g0(ap[k + 1], p);
        if (!st.g0Valid) {
            st.g0P24 = p[24];
            st.g0P25 = p[25];
            for (int k = 0; k < 6; ++k) st.g0Cache[k] = g[k];
            st.g0Valid = true;
        }
    }
    return (g[0] + (g[1] * ex + g[2] * ex * ex + g[3] * std::pow(ex, 3.0) +
                    (g[4] * std::pow(ex, 4.0) + g[5] * std::pow(ex, 12.0)) * (1.0 - std::pow(ex, 8.0)) / (1.0 - ex))) /
           sumex(ex);
}

// :639-686
inline void prepareGlobeBasis(State& st, const Input& in, const Flags& fl) {
    const double dgtr = 1.74533E-2, hr = 0.2618;
    const double tloc = in.lst;
    const double c = std::sin(in.g_lat * dgtr), s = std::cos(in.g_lat * dgtr);
    const double c2 = c * c, c4 = c2 * c2, s2 = s * s;
    double(&plg)[4][9] = st.plg;
    plg[0][1] = c;
    plg[0][2] = 0.5 * (3.0 * c2 - 1.0);
    plg[0][3] = 0.5 * (5.0 * c * c2 - 3.0 * c);
    plg[0][4] = (35.0 * c4 - 30.0 * c2 + 3.0) / 8.0;
    plg[0][5] = (63.0 * c2 * c2 * c - 70.0 * c2 * c + 15.0 * c) / 8.0;
    plg[0][6] = (11.0 * c * plg[0][5] - 5.0 * plg[0][4]) / 6.0;
    plg[1][1] = s;
    plg[1][2] = 3.0 * c * s;
    plg[1][3] = 1.5 * (5.0 * c2 - 1.0) * s;
    plg[1][4] = 2.5 * (7.0 * c2 * c - 3.0 * c) * s;
    plg[1][5] = 1.875 * (21.0 * c4 - 14.0 * c2 + 1.0) * s;
    plg[1][6] = (11.0 * c * plg[1][5] - 6.0 * plg[1][4]) / 5.0;
    plg[2][2] = 3.0 * s2;
    plg[2][3] = 15.0 * s2 * c;
    plg[2][4] = 7.5 * (7.0 * c2 - 1.0) * s2;
    plg[2][5] = 3.0 * c * plg[2][4] - 2.0 * plg[2][3];
    plg[2][6] = (11.0 * c * plg[2][5] - 7.0 * plg[2][4]) / 4.0;
    plg[2][7] = (13.0 * c * plg[2][6] - 8.0 * plg[2][5]) / 5.0;
    plg[3][3] = 15.0 * s2 * s;
    plg[3][4] = 105.0 * s2 * s * c;
    plg[3][5] = (9.0 * c * plg[3][4] - 7. * plg[3][3]) / 2.0;
    plg[3][6] = (11.0 * c * plg[3][5] - 8. * plg[3][4]) / 3.0;
    if (!(((fl.sw[7] == 0) && (fl.sw[8] == 0)) && (fl.sw[14] == 0))) {
        st.stloc = std::sin(hr * tloc);
        st.ctloc = std::cos(hr * tloc);
        st.s2tloc = std::sin(2.0 * hr * tloc);
        st.c2tloc = std::cos(2.0 * hr * tloc);
        st.s3tloc = std::sin(3.0 * hr * tloc);
        st.c3tloc = std::cos(3.0 * hr * tloc);
    }
}

// :688-850
inline double globe7(State& st, const double* p, const Input& in, const Flags& fl) {
    double t[15];
    const double sr = 7.2722E-5, dgtr = 1.74533E-2, dr = 1.72142E-2, hr = 0.2618;
    const double tloc = in.lst;
    const double(&plg)[4][9] = st.plg;
    for (int j = 0; j < 14; ++j) t[j] = 0;
    const double cd32 = std::cos(dr * (in.doy - p[31]));
    const double cd18 = std::cos(2.0 * dr * (in.doy - p[17]));
    const double cd14 = std::cos(dr * (in.doy - p[13]));
    const double cd39 = std::cos(2.0 * dr * (in.doy - p[38]));
    const double df = in.f107 - in.f107A;
    st.dfa = in.f107A - 150.0;
    const double dfa = st.dfa;
    t[0] = p[19] * df * (1.0 + p[59] * dfa) + p[20] * df * df + p[21] * dfa + p[29] * std::pow(dfa, 2.0);
    const double f1 = 1.0 + (p[47] * dfa + p[19] * df + p[20] * df * df) * fl.swc[1];
    const double f2 = 1.0 + (p[49] * dfa + p[19] * df + p[20] * df * df) * fl.swc[1];
    t[1] = (p[1] * plg[0][2] + p[2] * plg[0][4] + p[22] * plg[0][6]) + (p[14] * plg[0][2]) * dfa * fl.swc[1] + p[26] * plg[0][1];
    t[2] = p[18] * cd32;
    t[3] = (p[15] + p[16] * plg[0][2]) * cd18;
    t[4] = f1 * (p[9] * plg[0][1] + p[10] * plg[0][3]) * cd14;
    t[5] = p[37] * plg[0][1] * cd39;
    if (fl.sw[7]) {
        const double t71 = (p[11] * plg[1][2]) * cd14 * fl.swc[5];
        const double t72 = (p[12] * plg[1][2]) * cd14 * fl.swc[5];
        t[6] = f2 * ((p[3] * plg[1][1] + p[4] * plg[1][3] + p[27] * plg[1][5] + t71) * st.ctloc +
                     (p[6] * plg[1][1] + p[7] * plg[1][3] + p[28] * plg[1][5] + t72) * st.stloc);
    }
    if (fl.sw[8]) {
        const double t81 = (p[23] * plg[2][3] + p[35] * plg[2][5]) * cd14 * fl.swc[5];
        const double t82 = (p[33] * plg[2][3] + p[36] * plg[2][5]) * cd14 * fl.swc[5];
        t[7] = f2 * ((p[5] * plg[2][2] + p[41] * plg[2][4] + t81) * st.c2tloc + (p[8] * plg[2][2] + p[42] * plg[2][4] + t82) * st.s2tloc);
    }
    if (fl.sw[14]) {
        t[13] = f2 * ((p[39] * plg[3][3] + (p[93] * plg[3][4] + p[46] * plg[3][6]) * cd14 * fl.swc[5]) * st.s3tloc +
                      (p[40] * plg[3][3] + (p[94] * plg[3][4] + p[48] * plg[3][6]) * cd14 * fl.swc[5]) * st.c3tloc);
    }
    if (fl.sw[9] == -1) {
        if (p[51] != 0) {
            double exp1 = std::exp(-10800.0 * std::sqrt(p[51] * p[51]) / (1.0 + p[138] * (45.0 - std::sqrt(in.g_lat * in.g_lat))));
            if (exp1 > 0.99999) exp1 = 0.99999;
            st.apt[0] = sg0(st, exp1, p, in.apa);
            if (fl.sw[9]) {
                t[8] = st.apt[0] * (p[50] + p[96] * plg[0][2] + p[54] * plg[0][4] +
                                    (p[125] * plg[0][1] + p[126] * plg[0][3] + p[127] * plg[0][5]) * cd14 * fl.swc[5] +
                                    (p[128] * plg[1][1] + p[129] * plg[1][3] + p[130] * plg[1][5]) * fl.swc[7] *
                                        std::cos(hr * (tloc - p[131])));
            }
        }
    } else {
        const double apd = in.ap - 4.0;
        double p44 = p[43];
        const double p45 = p[44];
        if (p44 < 0) p44 = 1.0E-5;
        st.apdf = apd + (p45 - 1.0) * (apd + (std::exp(-p44 * apd) - 1.0) / p44);
        if (fl.sw[9]) {
            t[8] = st.apdf * (p[32] + p[45] * plg[0][2] + p[34] * plg[0][4] +
                              (p[100] * plg[0][1] + p[101] * plg[0][3] + p[102] * plg[0][5]) * cd14 * fl.swc[5] +
                              (p[121] * plg[1][1] + p[122] * plg[1][3] + p[123] * plg[1][5]) * fl.swc[7] *
                                  std::cos(hr * (tloc - p[124])));
        }
    }
    if ((fl.sw[10]) && (in.g_long > -1000.0)) {
        if (fl.sw[11]) {
            t[10] = (1.0 + p[80] * dfa * fl.swc[1]) *
                    ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[103] * plg[1][1] + p[104] * plg[1][3] +
                      p[105] * plg[1][5] + fl.swc[5] * (p[109] * plg[1][1] + p[110] * plg[1][3] + p[111] * plg[1][5]) * cd14) *
                         std::cos(dgtr * in.g_long) +
                     (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[106] * plg[1][1] + p[107] * plg[1][3] +
                      p[108] * plg[1][5] + fl.swc[5] * (p[112] * plg[1][1] + p[113] * plg[1][3] + p[114] * plg[1][5]) * cd14) *
                         std::sin(dgtr * in.g_long));
        }
        if (fl.sw[12]) {
            t[11] = (1.0 + p[95] * plg[0][1]) * (1.0 + p[81] * dfa * fl.swc[1]) * (1.0 + p[119] * plg[0][1] * fl.swc[5] * cd14) *
                    ((p[68] * plg[0][1] + p[69] * plg[0][3] + p[70] * plg[0][5]) * std::cos(sr * (in.sec - p[71])));
            t[11] += fl.swc[11] * (p[76] * plg[2][3] + p[77] * plg[2][5] + p[78] * plg[2][7]) *
                     std::cos(sr * (in.sec - p[79]) + 2.0 * dgtr * in.g_long) * (1.0 + p[137] * dfa * fl.swc[1]);
        }
        if (fl.sw[13]) {
            if (fl.sw[9] == -1) {
                if (p[51]) {
                    t[12] = st.apt[0] * fl.swc[11] * (1. + p[132] * plg[0][1]) *
                                ((p[52] * plg[1][2] + p[98] * plg[1][4] + p[67] * plg[1][6]) * std::cos(dgtr * (in.g_long - p[97]))) +
                            st.apt[0] * fl.swc[11] * fl.swc[5] * (p[133] * plg[1][1] + p[134] * plg[1][3] + p[135] * plg[1][5]) *
                                cd14 * std::cos(dgtr * (in.g_long - p[136])) +
                            st.apt[0] * fl.swc[12] * (p[55] * plg[0][1] + p[56] * plg[0][3] + p[57] * plg[0][5]) *
                                std::cos(sr * (in.sec - p[58]));
                }
            } else {
                t[12] = st.apdf * fl.swc[11] * (1.0 + p[120] * plg[0][1]) *
                            ((p[60] * plg[1][2] + p[61] * plg[1][4] + p[62] * plg[1][6]) * std::cos(dgtr * (in.g_long - p[63]))) +
                        st.apdf * fl.swc[11] * fl.swc[5] * (p[115] * plg[1][1] + p[116] * plg[1][3] + p[117] * plg[1][5]) * cd14 *
                            std::cos(dgtr * (in.g_long - p[118])) +
                        st.apdf * fl.swc[12] * (p[83] * plg[0][1] + p[84] * plg[0][3] + p[85] * plg[0][5]) *
                            std::cos(sr * (in.sec - p[75]));
            }
        }
    }
    double tinf = p[30];
    for (int i = 0; i < 14; ++i) tinf = tinf + std::fabs(double(fl.sw[i + 1])) * t[i];
    return tinf;
}

// :858-950
inline double glob7s(State& st, const double* p, const Input& in, const Flags& fl) {
    const double pset = 2.0;
    double t[14];
    const double dr = 1.72142E-2, dgtr = 1.74533E-2;
    const double(&plg)[4][9] = st.plg;
    const double parameterSet = (p[99] == 0.0) ? pset : p[99];
    if (parameterSet != pset) return -1;
    for (int j = 0; j < 14; ++j) t[j] = 0.0;
    const double cd32 = std::cos(dr * (in.doy - p[31]));
    const double cd18 = std::cos(2.0 * dr * (in.doy - p[17]));
    const double cd14 = std::cos(dr * (in.doy - p[13]));
    const double cd39 = std::cos(2.0 * dr * (in.doy - p[38]));
    t[0] = p[21] * st.dfa;
    t[1] = p[1] * plg[0][2] + p[2] * plg[0][4] + p[22] * plg[0][6] + p[26] * plg[0][1] + p[14] * plg[0][3] + p[59] * plg[0][5];
    t[2] = (p[18] + p[47] * plg[0][2] + p[29] * plg[0][4]) * cd32;
    t[3] = (p[15] + p[16] * plg[0][2] + p[30] * plg[0][4]) * cd18;
    t[4] = (p[9] * plg[0][1] + p[10] * plg[0][3] + p[20] * plg[0][5]) * cd14;
    t[5] = (p[37] * plg[0][1]) * cd39;
    if (fl.sw[7]) {
        const double t71 = p[11] * plg[1][2] * cd14 * fl.swc[5];
        const double t72 = p[12] * plg[1][2] * cd14 * fl.swc[5];
        t[6] = ((p[3] * plg[1][1] + p[4] * plg[1][3] + t71) * st.ctloc + (p[6] * plg[1][1] + p[7] * plg[1][3] + t72) * st.stloc);
    }
    if (fl.sw[8]) {
        const double t81 = (p[23] * plg[2][3] + p[35] * plg[2][5]) * cd14 * fl.swc[5];
        const double t82 = (p[33] * plg[2][3] + p[36] * plg[2][5]) * cd14 * fl.swc[5];
        t[7] = ((p[5] * plg[2][2] + p[41] * plg[2][4] + t81) * st.c2tloc + (p[8] * plg[2][2] + p[42] * plg[2][4] + t82) * st.s2tloc);
    }
    if (fl.sw[14]) t[13] = p[39] * plg[3][3] * st.s3tloc + p[40] * plg[3][3] * st.c3tloc;
    if (fl.sw[9]) {
        if (fl.sw[9] == 1) t[8] = st.apdf * (p[32] + p[45] * plg[0][2] * fl.swc[2]);
        if (fl.sw[9] == -1) t[8] = (p[50] * st.apt[0] + p[96] * plg[0][2] * st.apt[0] * fl.swc[2]);
    }
    if (!((fl.sw[10] == 0) || (fl.sw[11] == 0) || (in.g_long <= -1000.0))) {
        t[10] = (1.0 + plg[0][1] * (p[80] * fl.swc[5] * std::cos(dr * (in.doy - p[81])) + p[85] * fl.swc[6] * std::cos(2.0 * dr * (in.doy - p[86]))) +
                 p[83] * fl.swc[3] * std::cos(dr * (in.doy - p[84])) + p[87] * fl.swc[4] * std::cos(2.0 * dr * (in.doy - p[88]))) *
                ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[74] * plg[1][1] + p[75] * plg[1][3] + p[76] * plg[1][5]) *
                     std::cos(dgtr * in.g_long) +
                 (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[77] * plg[1][1] + p[78] * plg[1][3] + p[79] * plg[1][5]) *
                     std::sin(dgtr * in.g_long));
    }
    double tt = 0;
    for (int i = 0; i < 14; ++i) tt += std::fabs(double(fl.sw[i + 1])) * t[i];
    return tt;
}

// :1203-1498
inline void gts7(State& st, const NrlmsiseTables& T, Input& in, const Flags& fl, Output& out) {
    double zn1[5] = {120.0, 110.0, 100.0, 90.0, 72.5};
    const int mn1 = 5;
    const double dgtr = 1.74533E-2, dr = 1.72142E-2;
    const double alpha[9] = {-0.38, 0.0, 0.0, 0.0, 0.17, 0.0, -0.38, 0.0, 0.0};
    const double altl[8] = {200.0, 300.0, 160.0, 250.0, 240.0, 450.0, 320.0, 450.0};
    const double za = T.pdl[1][15];
    zn1[0] = za;
    for (int j = 0; j < 9; ++j) out.d[j] = 0;
    st.g0Valid = false;
    prepareGlobeBasis(st, in, fl);
    double tinf;
    if (in.alt > zn1[0]) tinf = T.ptm[0] * T.pt[0] * (1.0 + fl.sw[16] * globe7(st, T.pt, in, fl));
    else tinf = T.ptm[0] * T.pt[0];
    out.t[0] = tinf;
    double g0v;
    if (in.alt > zn1[4]) g0v = T.ptm[3] * T.ps[0] * (1.0 + fl.sw[19] * globe7(st, T.ps, in, fl));
    else g0v = T.ptm[3] * T.ps[0];
    const double tlb = T.ptm[1] * (1.0 + fl.sw[17] * globe7(st, T.pd[3], in, fl)) * T.pd[3][0];
    const double s = g0v / (tinf - tlb);
    if (in.alt < 300.0) {
        st.meso_tn1[1] = T.ptm[6] * T.ptl[0][0] / (1.0 - fl.sw[18] * glob7s(st, T.ptl[0], in, fl));
        st.meso_tn1[2] = T.ptm[2] * T.ptl[1][0] / (1.0 - fl.sw[18] * glob7s(st, T.ptl[1], in, fl));
        st.meso_tn1[3] = T.ptm[7] * T.ptl[2][0] / (1.0 - fl.sw[18] * glob7s(st, T.ptl[2], in, fl));
        st.meso_tn1[4] = T.ptm[4] * T.ptl[3][0] / (1.0 - fl.sw[18] * fl.sw[20] * glob7s(st, T.ptl[3], in, fl));
        st.meso_tgn1[1] = T.ptm[8] * T.pma[8][0] * (1.0 + fl.sw[18] * fl.sw[20] * glob7s(st, T.pma[8], in, fl)) * st.meso_tn1[4] *
                          st.meso_tn1[4] / (std::pow((T.ptm[4] * T.ptl[3][0]), 2.0));
    } else {
        st.meso_tn1[1] = T.ptm[6] * T.ptl[0][0];
        st.meso_tn1[2] = T.ptm[2] * T.ptl[1][0];
        st.meso_tn1[3] = T.ptm[7] * T.ptl[2][0];
        st.meso_tn1[4] = T.ptm[4] * T.ptl[3][0];
        st.meso_tgn1[1] = T.ptm[8] * T.pma[8][0] * st.meso_tn1[4] * st.meso_tn1[4] / (std::pow((T.ptm[4] * T.ptl[3][0]), 2.0));
    }
    const double g28 = fl.sw[21] * globe7(st, T.pd[2], in, fl);
    const double zhf = T.pdl[1][24] * (1.0 + fl.sw[5] * T.pdl[0][24] * std::sin(dgtr * in.g_lat) * std::cos(dr * (in.doy - T.pt[13])));
    out.t[0] = tinf;
    const double xmm = T.pdm[2][4];
    const double z = in.alt;
    double tz;
    double* tn1 = st.meso_tn1;
    double* tgn1 = st.meso_tgn1;

    // N2
    const double db28 = T.pdm[2][0] * std::exp(g28) * T.pd[2][0];
    out.d[2] = densu(st, z, db28, tinf, tlb, 28.0, alpha[2], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    const double zh28 = T.pdm[2][2] * zhf;
    const double zhm28 = T.pdm[2][3] * T.pdl[1][5];
    const double xmd = 28.0 - xmm;
    double b28 = 0.0;
    if (!(z > altl[7] && z >= zn1[0]))
        b28 = densu(st, zh28, db28, tinf, tlb, xmd, (alpha[2] - 1.0), &tz, T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[2])) {
        st.dm28 = densu(st, z, b28, tinf, tlb, xmm, alpha[2], &tz, T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[2] = dnet(out.d[2], st.dm28, zhm28, xmm, 28.0);
    }

    // He
    const double g4 = fl.sw[21] * globe7(st, T.pd[0], in, fl);
    const double db04 = T.pdm[0][0] * std::exp(g4) * T.pd[0][0];
    out.d[0] = densu(st, z, db04, tinf, tlb, 4., alpha[0], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z < altl[0])) {
        const double zh04 = T.pdm[0][2];
        const double b04 = densu(st, zh04, db04, tinf, tlb, 4. - xmm, alpha[0] - 1., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        st.dm04 = densu(st, z, b04, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[0] = dnet(out.d[0], st.dm04, zhm28, xmm, 4.);
        const double rl = std::log(b28 * T.pdm[0][1] / b04);
        const double zc04 = T.pdm[0][4] * T.pdl[1][0];
        const double hc04 = T.pdm[0][5] * T.pdl[1][1];
        out.d[0] = out.d[0] * ccor(z, rl, hc04, zc04);
    }

    // O
    const double g16 = fl.sw[21] * globe7(st, T.pd[1], in, fl);
    const double db16 = T.pdm[1][0] * std::exp(g16) * T.pd[1][0];
    out.d[1] = densu(st, z, db16, tinf, tlb, 16., alpha[1], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[1])) {
        const double zh16 = T.pdm[1][2];
        const double b16 = densu(st, zh16, db16, tinf, tlb, 16.0 - xmm, (alpha[1] - 1.0), &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        st.dm16 = densu(st, z, b16, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[1] = dnet(out.d[1], st.dm16, zhm28, xmm, 16.);
        const double rl = T.pdm[1][1] * T.pdl[1][16] * (1.0 + fl.sw[1] * T.pdl[0][23] * (in.f107A - 150.0));
        const double hc16 = T.pdm[1][5] * T.pdl[1][3];
        const double zc16 = T.pdm[1][4] * T.pdl[1][2];
        const double hc216 = T.pdm[1][5] * T.pdl[1][4];
        out.d[1] = out.d[1] * ccor2(z, rl, hc16, zc16, hc216);
        const double hcc16 = T.pdm[1][7] * T.pdl[1][13];
        const double zcc16 = T.pdm[1][6] * T.pdl[1][12];
        const double rc16 = T.pdm[1][3] * T.pdl[1][14];
        out.d[1] = out.d[1] * ccor(z, rc16, hcc16, zcc16);
    }

    // O2
    const double g32 = fl.sw[21] * globe7(st, T.pd[4], in, fl);
    const double db32 = T.pdm[3][0] * std::exp(g32) * T.pd[4][0];
    out.d[3] = densu(st, z, db32, tinf, tlb, 32., alpha[3], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if (fl.sw[15]) {
        if (z <= altl[3]) {
            const double zh32 = T.pdm[3][2];
            const double b32 = densu(st, zh32, db32, tinf, tlb, 32. - xmm, alpha[3] - 1., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
            st.dm32 = densu(st, z, b32, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
            out.d[3] = dnet(out.d[3], st.dm32, zhm28, xmm, 32.);
            const double rl = std::log(b28 * T.pdm[3][1] / b32);
            const double hc32 = T.pdm[3][5] * T.pdl[1][7];
            const double zc32 = T.pdm[3][4] * T.pdl[1][6];
            out.d[3] = out.d[3] * ccor(z, rl, hc32, zc32);
        }
        const double hcc32 = T.pdm[3][7] * T.pdl[1][22];
        const double hcc232 = T.pdm[3][7] * T.pdl[0][22];
        const double zcc32 = T.pdm[3][6] * T.pdl[1][21];
        const double rc32 = T.pdm[3][3] * T.pdl[1][23] * (1. + fl.sw[1] * T.pdl[0][23] * (in.f107A - 150.));
        out.d[3] = out.d[3] * ccor2(z, rc32, hcc32, zcc32, hcc232);
    }

    // Ar
    const double g40 = fl.sw[21] * globe7(st, T.pd[5], in, fl);
    const double db40 = T.pdm[4][0] * std::exp(g40) * T.pd[5][0];
    out.d[4] = densu(st, z, db40, tinf, tlb, 40., alpha[4], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[4])) {
        const double zh40 = T.pdm[4][2];
        const double b40 = densu(st, zh40, db40, tinf, tlb, 40. - xmm, alpha[4] - 1., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        st.dm40 = densu(st, z, b40, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[4] = dnet(out.d[4], st.dm40, zhm28, xmm, 40.);
        const double rl = std::log(b28 * T.pdm[4][1] / b40);
        const double hc40 = T.pdm[4][5] * T.pdl[1][9];
        const double zc40 = T.pdm[4][4] * T.pdl[1][8];
        out.d[4] = out.d[4] * ccor(z, rl, hc40, zc40);
    }

    // H
    const double g1 = fl.sw[21] * globe7(st, T.pd[6], in, fl);
    const double db01 = T.pdm[5][0] * std::exp(g1) * T.pd[6][0];
    out.d[6] = densu(st, z, db01, tinf, tlb, 1., alpha[6], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[6])) {
        const double zh01 = T.pdm[5][2];
        const double b01 = densu(st, zh01, db01, tinf, tlb, 1. - xmm, alpha[6] - 1., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        st.dm01 = densu(st, z, b01, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[6] = dnet(out.d[6], st.dm01, zhm28, xmm, 1.);
        const double rl = std::log(b28 * T.pdm[5][1] * std::sqrt(T.pdl[1][17] * T.pdl[1][17]) / b01);
        const double hc01 = T.pdm[5][5] * T.pdl[1][11];
        const double zc01 = T.pdm[5][4] * T.pdl[1][10];
        out.d[6] = out.d[6] * ccor(z, rl, hc01, zc01);
        const double hcc01 = T.pdm[5][7] * T.pdl[1][19];
        const double zcc01 = T.pdm[5][6] * T.pdl[1][18];
        const double rc01 = T.pdm[5][3] * T.pdl[1][20];
        out.d[6] = out.d[6] * ccor(z, rc01, hcc01, zcc01);
    }

    // N
    const double g14 = fl.sw[21] * globe7(st, T.pd[7], in, fl);
    const double db14 = T.pdm[6][0] * std::exp(g14) * T.pd[7][0];
    out.d[7] = densu(st, z, db14, tinf, tlb, 14., alpha[7], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[7])) {
        const double zh14 = T.pdm[6][2];
        const double b14 = densu(st, zh14, db14, tinf, tlb, 14. - xmm, alpha[7] - 1., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        st.dm14 = densu(st, z, b14, tinf, tlb, xmm, 0., &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
        out.d[7] = dnet(out.d[7], st.dm14, zhm28, xmm, 14.);
        const double rl = std::log(b28 * T.pdm[6][1] * std::sqrt(T.pdl[0][2] * T.pdl[0][2]) / b14);
        const double hc14 = T.pdm[6][5] * T.pdl[0][1];
        const double zc14 = T.pdm[6][4] * T.pdl[0][0];
        out.d[7] = out.d[7] * ccor(z, rl, hc14, zc14);
        const double hcc14 = T.pdm[6][7] * T.pdl[0][4];
        const double zcc14 = T.pdm[6][6] * T.pdl[0][3];
        const double rc14 = T.pdm[6][3] * T.pdl[0][5];
        out.d[7] = out.d[7] * ccor(z, rc14, hcc14, zcc14);
    }

    // anomalous O
    const double g16h = fl.sw[21] * globe7(st, T.pd[8], in, fl);
    const double db16h = T.pdm[7][0] * std::exp(g16h) * T.pd[8][0];
    const double tho = T.pdm[7][9] * T.pdl[0][6];
    const double ddo = densu(st, z, db16h, tho, tho, 16., alpha[8], &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    const double zsht = T.pdm[7][5];
    const double zmho = T.pdm[7][4];
    const double zsho = scalh(st, zmho, 16.0, tho);
    out.d[8] = ddo * std::exp(-zsht / zsho * (std::exp(-(z - zmho) / zsht) - 1.));

    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] + 14.0 * out.d[7]);
    const double zt = std::sqrt(in.alt * in.alt);
    (void)densu(st, zt, 1.0, tinf, tlb, 0.0, 0.0, &out.t[1], T.ptm[5], s, mn1, zn1, tn1, tgn1);
    if (fl.sw[0]) {
        for (int i = 0; i < 9; ++i) out.d[i] = out.d[i] * 1.0E6;
        out.d[5] = out.d[5] / 1000;
    }
}

// :958-1065
inline void gtd7(State& st, const NrlmsiseTables& T, Input& in, Flags& fl, Output& out) {
    const int mn3 = 5;
    const double zn3[5] = {32.5, 20.0, 15.0, 10.0, 0.0};
    const int mn2 = 4;
    const double zn2[4] = {72.5, 55.0, 45.0, 32.5};
    const double zmix = 62.5;
    Output so{};
    tselec(fl);
    double xlat = in.g_lat;
    if (fl.sw[2] == 0) xlat = 45.0;
    glatf(xlat, st.gsurf, st.re);
    const double xmm = T.pdm[2][4];
    double altt = in.alt > zn2[0] ? in.alt : zn2[0];
    const double tmp = in.alt;
    in.alt = altt;
    gts7(st, T, in, fl, so);
    altt = in.alt;
    in.alt = tmp;
    const double dm28m = fl.sw[0] ? st.dm28 * 1.0E6 : st.dm28;
    out.t[0] = so.t[0];
    out.t[1] = so.t[1];
    if (in.alt >= zn2[0]) {
        for (int i = 0; i < 9; ++i) out.d[i] = so.d[i];
        return;
    }
    st.meso_tgn2[0] = st.meso_tgn1[1];
    st.meso_tn2[0] = st.meso_tn1[4];
    st.meso_tn2[1] = T.pma[0][0] * T.pavgm[0] / (1.0 - fl.sw[20] * glob7s(st, T.pma[0], in, fl));
    st.meso_tn2[2] = T.pma[1][0] * T.pavgm[1] / (1.0 - fl.sw[20] * glob7s(st, T.pma[1], in, fl));
    st.meso_tn2[3] = T.pma[2][0] * T.pavgm[2] / (1.0 - fl.sw[20] * fl.sw[22] * glob7s(st, T.pma[2], in, fl));
    st.meso_tgn2[1] = T.pavgm[8] * T.pma[9][0] * (1.0 + fl.sw[20] * fl.sw[22] * glob7s(st, T.pma[9], in, fl)) * st.meso_tn2[3] *
                      st.meso_tn2[3] / (std::pow((T.pma[2][0] * T.pavgm[2]), 2.0));
    st.meso_tn3[0] = st.meso_tn2[3];
    if (in.alt <= zn3[0]) {
        st.meso_tgn3[0] = st.meso_tgn2[1];
        st.meso_tn3[1] = T.pma[3][0] * T.pavgm[3] / (1.0 - fl.sw[22] * glob7s(st, T.pma[3], in, fl));
        st.meso_tn3[2] = T.pma[4][0] * T.pavgm[4] / (1.0 - fl.sw[22] * glob7s(st, T.pma[4], in, fl));
        st.meso_tn3[3] = T.pma[5][0] * T.pavgm[5] / (1.0 - fl.sw[22] * glob7s(st, T.pma[5], in, fl));
        st.meso_tn3[4] = T.pma[6][0] * T.pavgm[6] / (1.0 - fl.sw[22] * glob7s(st, T.pma[6], in, fl));
        st.meso_tgn3[1] = T.pma[7][0] * T.pavgm[7] * (1.0 + fl.sw[22] * glob7s(st, T.pma[7], in, fl)) * st.meso_tn3[4] *
                          st.meso_tn3[4] / (std::pow((T.pma[6][0] * T.pavgm[6]), 2.0));
    }
    double dmc = 0;
    if (in.alt > zmix) dmc = 1.0 - (zn2[0] - in.alt) / (zn2[0] - zmix);
    const double dz28 = so.d[2];
    double tz;
    double dmr = so.d[2] / dm28m - 1.0;
    out.d[2] = densm(st, in.alt, dm28m, xmm, &tz, mn3, zn3, st.meso_tn3, st.meso_tgn3, mn2, zn2, st.meso_tn2, st.meso_tgn2);
    out.d[2] = out.d[2] * (1.0 + dmr * dmc);
    dmr = so.d[0] / (dz28 * T.pdm[0][1]) - 1.0;
    out.d[0] = out.d[2] * T.pdm[0][1] * (1.0 + dmr * dmc);
    out.d[1] = 0;
    out.d[8] = 0;
    dmr = so.d[3] / (dz28 * T.pdm[3][1]) - 1.0;
    out.d[3] = out.d[2] * T.pdm[3][1] * (1.0 + dmr * dmc);
    dmr = so.d[4] / (dz28 * T.pdm[4][1]) - 1.0;
    out.d[4] = out.d[2] * T.pdm[4][1] * (1.0 + dmr * dmc);
    out.d[6] = 0;
    out.d[7] = 0;
    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] + 14.0 * out.d[7]);
    if (fl.sw[0]) out.d[5] = out.d[5] / 1000;
    st.dd = densm(st, in.alt, 1.0, 0, &tz, mn3, zn3, st.meso_tn3, st.meso_tgn3, mn2, zn2, st.meso_tn2, st.meso_tgn2);
    out.t[1] = tz;
}

// :1073-1078
inline void gtd7d(State& st, const NrlmsiseTables& T, Input& in, Flags& fl, Output& out) {
    gtd7(st, T, in, fl, out);
    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] +
                           14.0 * out.d[7] + 16.0 * out.d[8]);
    if (fl.sw[0]) out.d[5] = out.d[5] / 1000;
}

// ---- input preparation: Nrlmsise00.cu:85-266 ----
constexpr double kSecondsPerDay = 86400.0;
constexpr double kMjdJ2000 = 51544.5;
constexpr double kRadToDeg = 57.2957795130823208768;
constexpr double kDegToRad = 0.0174532925199432957692;
constexpr double kTwoPi = 6.28318530717958647693;

// :91-100
inline double taiMinusUtc(double mjd) {
    double dat = 10.0;
    static const double dates[] = {41499, 41683, 42048, 42413, 42778, 43144, 43509, 43874, 44239,
                                   44786, 45151, 45516, 46247, 47161, 47892, 48257, 48804, 49169,
                                   49534, 50083, 50630, 51179, 53736, 54832, 56109, 57204, 57754};
    for (int i = 0; i < 27; ++i)
        if (mjd >= dates[i]) dat = 11.0 + i;
    return dat;
}

// :103-111
inline double etToMjdUtc(double et) {
    ORACLE_FLOP_SCOPE(kEpoch);
    const double days = et / kSecondsPerDay;
    const double g = 6.240075674 + 0.01720197034 * days;
    const double tdbMinusTt = 0.001657 * std::sin(g) + 0.000022 * std::sin(2.0 * g);
    const double mjdTt = kMjdJ2000 + (et - tdbMinusTt) / kSecondsPerDay;
    const double first = mjdTt - (32.184 + taiMinusUtc(mjdTt)) / kSecondsPerDay;
    return mjdTt - (32.184 + taiMinusUtc(first)) / kSecondsPerDay;
}

// :113-131
inline void calendarDate(double mjd, int& year, int& month, int& day, int& hour, int& minute, double& second) {
    ORACLE_FLOP_SCOPE(kEpoch);
    const long a = static_cast<long>(std::floor(mjd + 2400001.0));
    long b = 0, c = 0;
    if (a < 2299161) c = a + 1524;
    else {
        b = static_cast<long>((a - 1867216.25) / 36524.25);
        c = a + b - b / 4 + 1525;
    }
    const long d = static_cast<long>((c - 122.1) / 365.25);
    const long e = 365 * d + d / 4;
    const long f = static_cast<long>((c - e) / 30.6001);
    day = static_cast<int>(c - e - static_cast<long>(30.6001 * f));
    month = static_cast<int>(f - 1 - 12 * (f / 14));
    year = static_cast<int>(d - 4715 - (7 + month) / 10);
    double sod = (mjd - std::floor(mjd)) * kSecondsPerDay;
    if (sod < 0.0) sod += kSecondsPerDay;
    hour = static_cast<int>(sod / 3600.0);
    minute = static_cast<int>((sod - hour * 3600.0) / 60.0);
    second = sod - hour * 3600.0 - minute * 60.0;
}

// :133-140
inline int dayOfYear(int year, int month, int day) {
    int days[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if ((year % 4 == 0 && year % 100 != 0) || year % 400 == 0) days[1] = 29;
    int out = day;
    for (int m = 1; m < month; ++m) out += days[m - 1];
    return out;
}

// :142-146
inline double wrapHours(double hours) {
    hours = std::fmod(hours, 24.0);
    return hours < 0.0 ? hours + 24.0 : hours;
}

// :156-187
inline double hpopGastHours(double mjdUtc) {
    ORACLE_FLOP_SCOPE(kEpoch);
    const double mjdUt1 = mjdUtc;
    const double mjd0 = std::floor(mjdUt1);
    const double ut1 = kSecondsPerDay * (mjdUt1 - mjd0);
    const double t0 = (mjd0 - kMjdJ2000) / 36525.0;
    const double t = (mjdUt1 - kMjdJ2000) / 36525.0;
    const double gmstSeconds = 24110.54841 + 8640184.812866 * t0 + 1.002737909350795 * ut1 + (0.093104 - 6.2e-6 * t) * t * t;
    const double gmstRadians = kTwoPi * (gmstSeconds / kSecondsPerDay - std::floor(gmstSeconds / kSecondsPerDay));
    const double mjdTt = mjdUtc + (32.184 + taiMinusUtc(mjdUtc)) / kSecondsPerDay;
    const double tt = (mjdTt - kMjdJ2000) / 36525.0;
    const double omega = (125.04452 - 1934.136261 * tt) * kDegToRad;
    const double sunLongitude = (280.4665 + 36000.7698 * tt) * kDegToRad;
    const double moonLongitude = (218.3165 + 481267.8813 * tt) * kDegToRad;
    const double dpsiArcsec = -17.20 * std::sin(omega) - 1.32 * std::sin(2.0 * sunLongitude) - 0.23 * std::sin(2.0 * moonLongitude) +
                              0.21 * std::sin(2.0 * omega);
    const double meanObliquity = (23.43929111 - (46.8150 + (0.00059 - 0.001813 * tt) * tt) * tt / 3600.0) * kDegToRad;
    const double eqEq = dpsiArcsec / 3600.0 * kDegToRad * std::cos(meanObliquity);
    return wrapHours((gmstRadians + eqEq) * 24.0 / kTwoPi);
}

// :198-213
inline double weatherApAtLag(const NrlWeather& w, int record, int slot, int lag) {
    const int h = 24 + slot - lag;
    return w.at(5 + h % 8, record - 3 + h / 8);
}

// :215-241 (adjusted flux, the reference default)
inline bool fillWeather(const NrlWeather& w, double mjdUtc, double& f107, double& f107a, double& avgAp, double ap[7]) {
    ORACLE_FLOP_SCOPE(kEpoch);
    const int record = static_cast<int>(std::floor(mjdUtc - w.startMjdUtc));
    if (record < 3 || record >= w.days) return false;
    int slot = static_cast<int>(std::floor((mjdUtc - std::floor(mjdUtc)) * 8.0));
    if (slot < 0) slot = 0;
    if (slot > 7) slot = 7;
    f107 = w.at(2, record - 1);
    f107a = w.at(3, record);
    avgAp = w.at(4, record);
    ap[0] = avgAp;
    ap[1] = w.at(5 + slot, record);
    ap[2] = weatherApAtLag(w, record, slot, 1);
    ap[3] = weatherApAtLag(w, record, slot, 2);
    ap[4] = weatherApAtLag(w, record, slot, 3);
    ap[5] = ap[6] = 0.0;
    for (int d = 4; d <= 11; ++d) ap[5] += weatherApAtLag(w, record, slot, d);
    for (int d = 12; d <= 19; ++d) ap[6] += weatherApAtLag(w, record, slot, d);
    ap[5] /= 8.0;
    ap[6] /= 8.0;
    return true;
}

// :243-266
inline bool geodeticFixed(double x, double y, double z, double radius, double flattening, double& lat, double& lon, double& alt) {
    const double rho2 = x * x + y * y;
    if (!(rho2 + z * z > 0.0)) return false;
    const double e2 = flattening * (2.0 - flattening);
    double dz = e2 * z, zdz = z + dz, n = radius, nh = 0.0;
    for (int it = 0; it < 20; ++it) {
        zdz = z + dz;
        nh = std::sqrt(rho2 + zdz * zdz);
        const double sinPhi = zdz / nh;
        n = radius / std::sqrt(1.0 - e2 * sinPhi * sinPhi);
        const double next = n * e2 * sinPhi;
        if (std::fabs(dz - next) < 1.0e-12 * radius) {
            dz = next;
            break;
        }
        dz = next;
    }
    zdz = z + dz;
    nh = std::sqrt(rho2 + zdz * zdz);
    lon = std::atan2(y, x);
    lat = std::atan2(zdz, std::sqrt(rho2));
    alt = nh - n;
    return std::isfinite(lat) && std::isfinite(lon) && std::isfinite(alt);
}

}  // namespace nrl

// Nrlmsise00.cu:270-327 — density in kg/km^3 (0 above hCutoff, NaN on invalid data).
inline double nrlmsiseDensity(const NrlmsiseTables& T, const NrlWeather& w, double hCutoff, double densityScale,
                              const V3& fixed, double et, double Req, double flattening) {
    ORACLE_FLOP_SCOPE(kNrlmsise);
    using namespace nrl;
    const double mjdUtc = etToMjdUtc(et);
    double lat = 0, lon = 0, alt = 0;
    if (!geodeticFixed(fixed.x, fixed.y, fixed.z, Req, flattening, lat, lon, alt)) return std::nan("");
    if (alt > hCutoff) return 0.0;
    double f107 = 0, f107a = 0, avgAp = 0, ap[7] = {};
    if (!fillWeather(w, mjdUtc, f107, f107a, avgAp, ap)) return std::nan("");
    int year = 0, month = 0, day = 0, hour = 0, minute = 0;
    double second = 0;
    calendarDate(mjdUtc, year, month, day, hour, minute, second);
    const double sec = hour * 3600.0 + minute * 60.0 + second;
    const double lst = wrapHours(hpopGastHours(mjdUtc) + lon * kRadToDeg / 15.0);
    Input in{};
    in.doy = dayOfYear(year, month, day);
    in.sec = sec;
    in.alt = alt;
    in.g_lat = lat * kRadToDeg;
    in.g_long = lon * kRadToDeg;
    in.lst = lst;
    in.f107A = f107a;
    in.f107 = f107;
    in.ap = avgAp;
    for (int k = 0; k < 7; ++k) in.apa[k] = ap[k];
    Flags fl{};
    for (int k = 0; k < 24; ++k) fl.switches[k] = 1;
    fl.switches[9] = -1;
    Output out{};
    State st;
    gtd7d(st, T, in, fl, out);
    if (!std::isfinite(out.d[5]) || out.d[5] < 0.0) return std::nan("");
    return (out.d[5] * 1.0e9) * densityScale;
}

// Raw model call for unit tests: returns total mass density d[5] (kg/m^3 with
// switch 0 on) and the two temperatures.
inline void nrlmsiseRaw(const NrlmsiseTables& T, int doy, double sec, double alt, double glat, double glon, double lst,
                        double f107a, double f107, double ap, const double* apa, int useApArray, double out9[9], double t2[2]) {
    using namespace nrl;
    Input in{};
    in.doy = doy;
    in.sec = sec;
    in.alt = alt;
    in.g_lat = glat;
    in.g_long = glon;
    in.lst = lst;
    in.f107A = f107a;
    in.f107 = f107;
    in.ap = ap;
    for (int k = 0; k < 7; ++k) in.apa[k] = apa ? apa[k] : ap;
    Flags fl{};
    for (int k = 0; k < 24; ++k) fl.switches[k] = 1;
    fl.switches[0] = 0;  // cgs, as in the published test driver
    if (useApArray) fl.switches[9] = -1;
    Output out{};
    State st;
    gtd7d(st, T, in, fl, out);
    for (int k = 0; k < 9; ++k) out9[k] = out.d[k];
    t2[0] = out.t[0];
    t2[1] = out.t[1];
}

}  // namespace oracle
