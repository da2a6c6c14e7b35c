// NRLMSISE-00 total-density model on the device (one object per lane).
// Follows the reference model file
// src/common/engine/paraHPOP/model/environment/atmosphere/detail/Nrlmsise00Model.inl
// (tselec :80-100, glatf :108-114, ccor/ccor2 :120-164, scalh :172-178,
// dnet :186-215, splini/splint/spline :223-338, zeta/densm :346-468,
// densu :476-576, g0/sumex/sg0 :586-634, prepareGlobeBasis :639-686,
// globe7 :688-850, glob7s :858-950, gtd7 :958-1065, gtd7d :1073-1078,
// gts7 :1203-1498). The model's file-scope variables live in a per-lane State.
// Coefficient blocks come from one device array (offsets below), read with
// uniform addresses.
#pragma once

#include <hip/hip_runtime.h>

#include "phx_common.h"

namespace phx {
namespace nrlmsise {

// Coefficient tables and per-epoch records are read-only during a kernel: the
// fast paths read them through constant-address-space pointers, so uniform
// indices become scalar loads into SGPR operands.
typedef const __attribute__((address_space(4))) double* NrlCD;

enum : int {
    OFF_PT = 0, OFF_PD = 150, OFF_PS = 1500, OFF_PDL = 1650, OFF_PTM = 1700, OFF_PDM = 1710,
    OFF_PTL = 1790, OFF_PMA = 2190, OFF_SAM = 3190, OFF_PAVGM = 3290, COEFF_COUNT = 3300
};

// Coefficients that gts7's prefix, the shared densu quantities and the species
// read one at a time on their dependency chains, gathered after the table
// (COEFF_COUNT values, then these): pdl, ptm and pdm whole, pt[0..14], ps[0] and
// the first entries of the pd, ptl and pma[8] blocks.
enum : int {
    HOT_PDL = 0,      // pdl[2][25]
    HOT_PTM = 50,     // ptm[10]
    HOT_PDM = 60,     // pdm[8][10]
    HOT_PT = 140,     // pt[0..14]
    HOT_PS0 = 155,    // ps[0]
    HOT_PD0 = 156,    // pd[k][0], k = 0..8
    HOT_PTL0 = 165,   // ptl[k][0], k = 0..3
    HOT_PMA80 = 169,  // pma[8][0]
    HOT_COUNT = 170
};

// `hot` is the gathered block: behind the table in device memory, or a copy in
// LDS (the drag tile kernel's), read on the chains instead of scalar loads.
struct NrlTables {
    const double* base;
    const double* hot;
    __device__ const double* pt() const { return base + OFF_PT; }
};

__device__ inline NrlTables nrlTables(const double* coeff) { return NrlTables{coeff, coeff + COEFF_COUNT}; }



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
    double cglong = 0, sglong = 0;   // cos/sin(dgtr * g_long), fast path
    bool g0Valid = false;
    double g0P24 = 0, g0P25 = 0, g0Cache[6] = {};
};

// Nrlmsise00Model.inl:80-100
__device__ inline void tselec(Flags& f) {
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
__device__ inline void glatf(double lat, double& gv, double& reff) {
    const double dgtr = 1.74533E-2;
    const double c2 = cos(2.0 * dgtr * lat);
    gv = 980.616 * (1.0 - 0.0026373 * c2);
    reff = 2.0 * gv / (3.085462E-6 + 2.27E-9 * c2) * 1.0E-5;
}

// :120-137
__device__ inline double ccor(double alt, double r, double h1, double zh) {
    double e = (alt - zh) / h1;
    if (e > 70) return exp(0.0);
    if (e < -70) return exp(r);
    const double ex = exp(e);
    e = r / (1.0 + ex);
    return exp(e);
}

// ccor split at its argument r: the part that depends on the altitude alone
// (the branch taken and 1 + exp((alt - zh) / h1)), then the factor for r;
// together the same operations as ccor.
struct CcorPre {
    int mode;     // 0: exp(0), 1: exp(r), 2: exp(r / den)
    double den;
};
__device__ inline CcorPre ccorPre(double alt, double h1, double zh) {
    const double e = (alt - zh) / h1;
    if (e > 70) return {0, 0.0};
    if (e < -70) return {1, 0.0};
    return {2, 1.0 + exp(e)};
}
__device__ inline double ccorFrom(const CcorPre& c, double r) {
    if (c.mode == 0) return exp(0.0);
    if (c.mode == 1) return exp(r);
    return exp(r / c.den);
}

// :145-164
__device__ inline double ccor2(double alt, double r, double h1, double zh, double h2) {
    const double e1 = (alt - zh) / h1, e2 = (alt - zh) / h2;
    if (e1 > 70 || e2 > 70) return exp(0.0);
    if (e1 < -70 && e2 < -70) return exp(r);
    const double ex1 = exp(e1), ex2 = exp(e2);
    return exp(r / (1.0 + 0.5 * (ex1 + ex2)));
}

// :172-178
__device__ inline double scalh(const State& st, double alt, double xm, double temp) {
    const double rgas = 831.4;
    double g = st.gsurf / pow((1.0 + alt / st.re), 2.0);
    g = rgas * temp / (g * xm);
    return g;
}

// :186-215
__device__ inline double dnet(double dd, double dm, double zhm, double xmm, double xm) {
    double a = zhm / (xmm - xm);
    if (!((dm > 0) && (dd > 0))) {
        if (dd == 0 && dm == 0) dd = 1;
        if (dm == 0) return dd;
        if (dd == 0) return dm;
    }
    const double ylog = a * log(dm / dd);
    if (ylog < -10) return dd;
    if (ylog > 10) return dm;
    a = dd * pow((1.0 + exp(ylog)), (1.0 / a));
    return a;
}

// dnet with (1 + e^ylog)^(1/a) taken as exp(log(1 + e^ylog) / a) (species fast path)
__device__ inline double dnetFast(double dd, double dm, double zhm, double xmm, double xm) {
    double a = zhm / (xmm - xm);
    if (!((dm > 0) && (dd > 0))) {
        if (dd == 0 && dm == 0) dd = 1;
        if (dm == 0) return dd;
        if (dd == 0) return dm;
    }
    const double ylog = a * log(dm / dd);
    if (ylog < -10) return dd;
    if (ylog > 10) return dm;
    return dd * exp(log(1.0 + exp(ylog)) * (1.0 / a));
}

// :223-252
__device__ inline void splini(const double* xa, const double* ya, const double* y2a, int n, double x, double* y) {
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
__device__ inline void splint(const double* xa, const double* ya, const double* y2a, int n, double x, double* y) {
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
__device__ inline void spline(const double* x, const double* y, int n, double yp1, double ypn, double* y2) {
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
// Fixed-size forms of the three spline routines above for the N = 5 node
// profiles of densu: every array index is a compile-time constant after
// unrolling, so the nodes stay in registers (the pointer forms index private
// memory). Same arithmetic in the same order; the interval search of splint is
// the binary search's result for increasing nodes (the zeta-scaled nodes of
// densu increase strictly): the last node at or below x, clamped to [0, N-2].
template <int N>
__device__ inline void spliniN(const double (&xa)[N], const double (&ya)[N], const double (&y2a)[N], double x, double* y) {
    double yi = 0;
    bool go = true;
#pragma unroll
    for (int klo = 0; klo < N - 1; ++klo) {
        const int khi = klo + 1;
        go = go && (x > xa[klo]);
        if (go) {
            double xx = x;
            if (khi < (N - 1)) xx = x < xa[khi] ? x : xa[khi];
            const double h = xa[khi] - xa[klo];
            const double a = (xa[khi] - xx) / h;
            const double b = (xx - xa[klo]) / h;
            const double a2 = a * a, b2 = b * b;
            yi += ((1.0 - a2) * ya[klo] / 2.0 + b2 * ya[khi] / 2.0 +
                   ((-(1.0 + a2 * a2) / 4.0 + a2 / 2.0) * y2a[klo] + (b2 * b2 / 4.0 - b2 / 2.0) * y2a[khi]) * h * h / 6.0) * h;
        }
    }
    *y = yi;
}

template <int N>
__device__ inline void splintN(const double (&xa)[N], const double (&ya)[N], const double (&y2a)[N], double x, double* y) {
    // interval values carried by conditional moves (no index variable)
    double xlo = xa[0], xhi = xa[1], ylo = ya[0], yhi = ya[1], y2lo = y2a[0], y2hi = y2a[1];
#pragma unroll
    for (int k = 1; k < N - 1; ++k) {
        const bool up = xa[k] <= x;
        xlo = up ? xa[k] : xlo;
        xhi = up ? xa[k + 1] : xhi;
        ylo = up ? ya[k] : ylo;
        yhi = up ? ya[k + 1] : yhi;
        y2lo = up ? y2a[k] : y2lo;
        y2hi = up ? y2a[k + 1] : y2hi;
    }
    const double h = xhi - xlo;
    const double a = (xhi - x) / h;
    const double b = (x - xlo) / h;
    *y = a * ylo + b * yhi + ((a * a * a - a) * y2lo + (b * b * b - b) * y2hi) * h * h / 6.0;
}

template <int N>
__device__ inline void splineN(const double (&x)[N], const double (&y)[N], double yp1, double ypn, double (&y2)[N]) {
    double u[N];
    if (yp1 > 0.99E30) {
        y2[0] = 0;
        u[0] = 0;
    } else {
        y2[0] = -0.5;
        u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
#pragma unroll
    for (int i = 1; i < (N - 1); ++i) {
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
        un = (3.0 / (x[N - 1] - x[N - 2])) * (ypn - (y[N - 1] - y[N - 2]) / (x[N - 1] - x[N - 2]));
    }
    y2[N - 1] = (un - qn * u[N - 2]) / (qn * y2[N - 2] + 1.0);
#pragma unroll
    for (int k = N - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

__device__ inline double zeta(const State& st, double zz, double zl) { return ((zz - zl) * (st.re + zl) / (st.re + zz)); }

// :350-468
__device__ inline double densm(State& st, double alt, double d0, double xm, double* tz, int mn3, const double* zn3, const double* tn3,
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
    double yd2 = -tgn2[1] / (t2 * t2) * zgdif * (pow(((st.re + z2) / (st.re + z1)), 2.0));
    spline(xs, ys, mn, yd1, yd2, y2out);
    double x = zg / zgdif, y;
    splint(xs, ys, y2out, mn, x, &y);
    *tz = 1.0 / y;
    if (xm != 0.0) {
        const double glb = st.gsurf / (pow((1.0 + z1 / st.re), 2.0));
        const double gamm = xm * glb * zgdif / rgas;
        double yi;
        splini(xs, ys, y2out, mn, x, &yi);
        double expl = gamm * yi;
        if (expl > 50.0) expl = 50.0;
        densm_tmp = densm_tmp * (t1 / *tz) * exp(-expl);
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
    yd2 = -tgn3[1] / (t2 * t2) * zgdif * (pow(((st.re + z2) / (st.re + z1)), 2.0));
    spline(xs, ys, mn, yd1, yd2, y2out);
    x = zg / zgdif;
    splint(xs, ys, y2out, mn, x, &y);
    *tz = 1.0 / y;
    if (xm != 0.0) {
        const double glb = st.gsurf / (pow((1.0 + z1 / st.re), 2.0));
        const double gamm = xm * glb * zgdif / rgas;
        double yi;
        splini(xs, ys, y2out, mn, x, &yi);
        double expl = gamm * yi;
        if (expl > 50.0) expl = 50.0;
        densm_tmp = densm_tmp * (t1 / *tz) * exp(-expl);
    }
    return xm == 0.0 ? *tz : densm_tmp;
}

// :476-576
// mn1 is 5 at every call site (the five lower-thermosphere nodes of gts7).
__device__ inline double densu(State& st, double alt, double dlb, double tinf, double tlb, double xm, double alpha, double* tz,
                    double zlb, double s2, int mn1, const double* zn1, double* tn1, double* tgn1) {
    constexpr int MN = 5;
    (void)mn1;
    double x = 0, y;
    const double rgas = 831.4;
    double densu_temp = 1.0;
    double z1 = 0, t1 = 0, zgdif = 0;
    double xs[MN] = {}, ys[MN] = {}, y2out[MN] = {};
    const double za = zn1[0];
    double z = alt > za ? alt : za;
    const double zg2 = zeta(st, z, zlb);
    const double tt = tinf - (tinf - tlb) * exp(-s2 * zg2);
    const double ta = tt;
    *tz = tt;
    densu_temp = *tz;
    if (alt < za) {
        const double dta = (tinf - ta) * s2 * pow(((st.re + zlb) / (st.re + za)), 2.0);
        tgn1[0] = dta;
        tn1[0] = ta;
        z = alt > zn1[MN - 1] ? alt : zn1[MN - 1];
        z1 = zn1[0];
        const double z2 = zn1[MN - 1];
        t1 = tn1[0];
        const double t2 = tn1[MN - 1];
        const double zg = zeta(st, z, z1);
        zgdif = zeta(st, z2, z1);
#pragma unroll
        for (int k = 0; k < MN; ++k) {
            xs[k] = zeta(st, zn1[k], z1) / zgdif;
            ys[k] = 1.0 / tn1[k];
        }
        const double yd1 = -tgn1[0] / (t1 * t1) * zgdif;
        const double yd2 = -tgn1[1] / (t2 * t2) * zgdif * pow(((st.re + z2) / (st.re + z1)), 2.0);
        splineN(xs, ys, yd1, yd2, y2out);
        x = zg / zgdif;
        splintN(xs, ys, y2out, x, &y);
        *tz = 1.0 / y;
        densu_temp = *tz;
    }
    if (xm == 0) return densu_temp;
    double glb = st.gsurf / pow((1.0 + zlb / st.re), 2.0);
    const double gamma = xm * glb / (s2 * rgas * tinf);
    double expl = exp(-s2 * gamma * zg2);
    if (expl > 50.0) expl = 50.0;
    if (tt <= 0) expl = 50.0;
    const double densa = dlb * pow((tlb / tt), ((1.0 + alpha + gamma))) * expl;
    densu_temp = densa;
    if (alt >= za) return densu_temp;
    glb = st.gsurf / pow((1.0 + z1 / st.re), 2.0);
    const double gamm = xm * glb * zgdif / rgas;
    double yi;
    spliniN(xs, ys, y2out, x, &yi);
    expl = gamm * yi;
    if (expl > 50.0) expl = 50.0;
    if (*tz <= 0) expl = 50.0;
    densu_temp = densu_temp * pow((t1 / *tz), (1.0 + alpha)) * exp(-expl);
    return densu_temp;
}

// :586-593 (p[24] lower bound applied to a local, as in the reference)
__device__ inline double g0(double a, const double* p) {
    const double p24 = (p[24] < 1.0E-4) ? 1.0E-4 : sqrt(p[24] * p[24]);
    return (a - 4.0 + (p[25] - 1.0) * (a - 4.0 + (exp(-p24 * (a - 4.0)) - 1.0) / p24));
}

// :596-598
__device__ inline double sumex(double ex) { return (1.0 + (1.0 - pow(ex, 19.0)) / (1.0 - ex) * pow(ex, 0.5)); }

// :601-634
__device__ inline double sg0(State& st, double ex, const double* p, const double* ap) {
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
    return (g[0] + (g[1] * ex + g[2] * ex * ex + g[3] * pow(ex, 3.0) +
                    (g[4] * pow(ex, 4.0) + g[5] * pow(ex, 12.0)) * (1.0 - pow(ex, 8.0)) / (1.0 - ex))) /
           sumex(ex);
}

// :639-686
__device__ inline void prepareGlobeBasis(State& st, const Input& in, const Flags& fl) {
    const double dgtr = 1.74533E-2, hr = 0.2618;
    const double tloc = in.lst;
    const double c = sin(in.g_lat * dgtr), s = cos(in.g_lat * dgtr);
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
        st.stloc = sin(hr * tloc);
        st.ctloc = cos(hr * tloc);
        st.s2tloc = sin(2.0 * hr * tloc);
        st.c2tloc = cos(2.0 * hr * tloc);
        st.s3tloc = sin(3.0 * hr * tloc);
        st.c3tloc = cos(3.0 * hr * tloc);
    }
}

// :688-850
__device__ inline double globe7(State& st, const double* p, const Input& in, const Flags& fl) {
    double t[15];
    const double sr = 7.2722E-5, dgtr = 1.74533E-2, dr = 1.72142E-2, hr = 0.2618;
    const double tloc = in.lst;
    const double(&plg)[4][9] = st.plg;
    for (int j = 0; j < 14; ++j) t[j] = 0;
    const double cd32 = cos(dr * (in.doy - p[31]));
    const double cd18 = cos(2.0 * dr * (in.doy - p[17]));
    const double cd14 = cos(dr * (in.doy - p[13]));
    const double cd39 = cos(2.0 * dr * (in.doy - p[38]));
    const double df = in.f107 - in.f107A;
    st.dfa = in.f107A - 150.0;
    const double dfa = st.dfa;
    t[0] = p[19] * df * (1.0 + p[59] * dfa) + p[20] * df * df + p[21] * dfa + p[29] * pow(dfa, 2.0);
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
            double exp1 = exp(-10800.0 * sqrt(p[51] * p[51]) / (1.0 + p[138] * (45.0 - sqrt(in.g_lat * in.g_lat))));
            if (exp1 > 0.99999) exp1 = 0.99999;
            st.apt[0] = sg0(st, exp1, p, in.apa);
            if (fl.sw[9]) {
                t[8] = st.apt[0] * (p[50] + p[96] * plg[0][2] + p[54] * plg[0][4] +
                                    (p[125] * plg[0][1] + p[126] * plg[0][3] + p[127] * plg[0][5]) * cd14 * fl.swc[5] +
                                    (p[128] * plg[1][1] + p[129] * plg[1][3] + p[130] * plg[1][5]) * fl.swc[7] *
                                        cos(hr * (tloc - p[131])));
            }
        }
    } else {
        const double apd = in.ap - 4.0;
        double p44 = p[43];
        const double p45 = p[44];
        if (p44 < 0) p44 = 1.0E-5;
        st.apdf = apd + (p45 - 1.0) * (apd + (exp(-p44 * apd) - 1.0) / p44);
        if (fl.sw[9]) {
            t[8] = st.apdf * (p[32] + p[45] * plg[0][2] + p[34] * plg[0][4] +
                              (p[100] * plg[0][1] + p[101] * plg[0][3] + p[102] * plg[0][5]) * cd14 * fl.swc[5] +
                              (p[121] * plg[1][1] + p[122] * plg[1][3] + p[123] * plg[1][5]) * fl.swc[7] *
                                  cos(hr * (tloc - p[124])));
        }
    }
    if ((fl.sw[10]) && (in.g_long > -1000.0)) {
        if (fl.sw[11]) {
            t[10] = (1.0 + p[80] * dfa * fl.swc[1]) *
                    ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[103] * plg[1][1] + p[104] * plg[1][3] +
                      p[105] * plg[1][5] + fl.swc[5] * (p[109] * plg[1][1] + p[110] * plg[1][3] + p[111] * plg[1][5]) * cd14) *
                         cos(dgtr * in.g_long) +
                     (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[106] * plg[1][1] + p[107] * plg[1][3] +
                      p[108] * plg[1][5] + fl.swc[5] * (p[112] * plg[1][1] + p[113] * plg[1][3] + p[114] * plg[1][5]) * cd14) *
                         sin(dgtr * in.g_long));
        }
        if (fl.sw[12]) {
            t[11] = (1.0 + p[95] * plg[0][1]) * (1.0 + p[81] * dfa * fl.swc[1]) * (1.0 + p[119] * plg[0][1] * fl.swc[5] * cd14) *
                    ((p[68] * plg[0][1] + p[69] * plg[0][3] + p[70] * plg[0][5]) * cos(sr * (in.sec - p[71])));
            t[11] += fl.swc[11] * (p[76] * plg[2][3] + p[77] * plg[2][5] + p[78] * plg[2][7]) *
                     cos(sr * (in.sec - p[79]) + 2.0 * dgtr * in.g_long) * (1.0 + p[137] * dfa * fl.swc[1]);
        }
        if (fl.sw[13]) {
            if (fl.sw[9] == -1) {
                if (p[51]) {
                    t[12] = st.apt[0] * fl.swc[11] * (1. + p[132] * plg[0][1]) *
                                ((p[52] * plg[1][2] + p[98] * plg[1][4] + p[67] * plg[1][6]) * cos(dgtr * (in.g_long - p[97]))) +
                            st.apt[0] * fl.swc[11] * fl.swc[5] * (p[133] * plg[1][1] + p[134] * plg[1][3] + p[135] * plg[1][5]) *
                                cd14 * cos(dgtr * (in.g_long - p[136])) +
                            st.apt[0] * fl.swc[12] * (p[55] * plg[0][1] + p[56] * plg[0][3] + p[57] * plg[0][5]) *
                                cos(sr * (in.sec - p[58]));
                }
            } else {
                t[12] = st.apdf * fl.swc[11] * (1.0 + p[120] * plg[0][1]) *
                            ((p[60] * plg[1][2] + p[61] * plg[1][4] + p[62] * plg[1][6]) * cos(dgtr * (in.g_long - p[63]))) +
                        st.apdf * fl.swc[11] * fl.swc[5] * (p[115] * plg[1][1] + p[116] * plg[1][3] + p[117] * plg[1][5]) * cd14 *
                            cos(dgtr * (in.g_long - p[118])) +
                        st.apdf * fl.swc[12] * (p[83] * plg[0][1] + p[84] * plg[0][3] + p[85] * plg[0][5]) *
                            cos(sr * (in.sec - p[75]));
            }
        }
    }
    double tinf = p[30];
    for (int i = 0; i < 14; ++i) tinf = tinf + fabs(double(fl.sw[i + 1])) * t[i];
    return tinf;
}

// :858-950
__device__ inline double glob7s(State& st, const double* p, const Input& in, const Flags& fl) {
    const double pset = 2.0;
    double t[14];
    const double dr = 1.72142E-2, dgtr = 1.74533E-2;
    const double(&plg)[4][9] = st.plg;
    const double parameterSet = (p[99] == 0.0) ? pset : p[99];
    if (parameterSet != pset) return -1;
    for (int j = 0; j < 14; ++j) t[j] = 0.0;
    const double cd32 = cos(dr * (in.doy - p[31]));
    const double cd18 = cos(2.0 * dr * (in.doy - p[17]));
    const double cd14 = cos(dr * (in.doy - p[13]));
    const double cd39 = cos(2.0 * dr * (in.doy - p[38]));
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
        t[10] = (1.0 + plg[0][1] * (p[80] * fl.swc[5] * cos(dr * (in.doy - p[81])) + p[85] * fl.swc[6] * cos(2.0 * dr * (in.doy - p[86]))) +
                 p[83] * fl.swc[3] * cos(dr * (in.doy - p[84])) + p[87] * fl.swc[4] * cos(2.0 * dr * (in.doy - p[88]))) *
                ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[74] * plg[1][1] + p[75] * plg[1][3] + p[76] * plg[1][5]) *
                     cos(dgtr * in.g_long) +
                 (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[77] * plg[1][1] + p[78] * plg[1][3] + p[79] * plg[1][5]) *
                     sin(dgtr * in.g_long));
    }
    double tt = 0;
    for (int i = 0; i < 14; ++i) tt += fabs(double(fl.sw[i + 1])) * t[i];
    return tt;
}

// :1203-1498
template <class Globe>
__device__ inline void gts7(State& st, const NrlTables& T, Input& in, const Flags& fl, Output& out, const Globe& G) {
    double zn1[5] = {120.0, 110.0, 100.0, 90.0, 72.5};
    const int mn1 = 5;
    const double dgtr = 1.74533E-2, dr = 1.72142E-2;
    const double alpha[9] = {-0.38, 0.0, 0.0, 0.0, 0.17, 0.0, -0.38, 0.0, 0.0};
    const double altl[8] = {200.0, 300.0, 160.0, 250.0, 240.0, 450.0, 320.0, 450.0};
    const double za = (T.base+OFF_PDL+25*(1))[15];
    zn1[0] = za;
    for (int j = 0; j < 9; ++j) out.d[j] = 0;
    st.g0Valid = false;
    G.basis(st, in, fl);
    double tinf;
    if (in.alt > zn1[0]) tinf = (T.base+OFF_PTM)[0] * (T.base+OFF_PT)[0] * (1.0 + fl.sw[16] * G(st, 0, (T.base+OFF_PT), in, fl));
    else tinf = (T.base+OFF_PTM)[0] * (T.base+OFF_PT)[0];
    out.t[0] = tinf;
    double g0v;
    if (in.alt > zn1[4]) g0v = (T.base+OFF_PTM)[3] * (T.base+OFF_PS)[0] * (1.0 + fl.sw[19] * G(st, 1, (T.base+OFF_PS), in, fl));
    else g0v = (T.base+OFF_PTM)[3] * (T.base+OFF_PS)[0];
    const double tlb = (T.base+OFF_PTM)[1] * (1.0 + fl.sw[17] * G(st, 5, (T.base+OFF_PD+150*(3)), in, fl)) * (T.base+OFF_PD+150*(3))[0];
    const double s = g0v / (tinf - tlb);
    if (in.alt < 300.0) {
        st.meso_tn1[1] = (T.base+OFF_PTM)[6] * (T.base+OFF_PTL+100*(0))[0] / (1.0 - fl.sw[18] * glob7s(st, (T.base+OFF_PTL+100*(0)), in, fl));
        st.meso_tn1[2] = (T.base+OFF_PTM)[2] * (T.base+OFF_PTL+100*(1))[0] / (1.0 - fl.sw[18] * glob7s(st, (T.base+OFF_PTL+100*(1)), in, fl));
        st.meso_tn1[3] = (T.base+OFF_PTM)[7] * (T.base+OFF_PTL+100*(2))[0] / (1.0 - fl.sw[18] * glob7s(st, (T.base+OFF_PTL+100*(2)), in, fl));
        st.meso_tn1[4] = (T.base+OFF_PTM)[4] * (T.base+OFF_PTL+100*(3))[0] / (1.0 - fl.sw[18] * fl.sw[20] * glob7s(st, (T.base+OFF_PTL+100*(3)), in, fl));
        st.meso_tgn1[1] = (T.base+OFF_PTM)[8] * (T.base+OFF_PMA+100*(8))[0] * (1.0 + fl.sw[18] * fl.sw[20] * glob7s(st, (T.base+OFF_PMA+100*(8)), in, fl)) * st.meso_tn1[4] *
                          st.meso_tn1[4] / (pow(((T.base+OFF_PTM)[4] * (T.base+OFF_PTL+100*(3))[0]), 2.0));
    } else {
        st.meso_tn1[1] = (T.base+OFF_PTM)[6] * (T.base+OFF_PTL+100*(0))[0];
        st.meso_tn1[2] = (T.base+OFF_PTM)[2] * (T.base+OFF_PTL+100*(1))[0];
        st.meso_tn1[3] = (T.base+OFF_PTM)[7] * (T.base+OFF_PTL+100*(2))[0];
        st.meso_tn1[4] = (T.base+OFF_PTM)[4] * (T.base+OFF_PTL+100*(3))[0];
        st.meso_tgn1[1] = (T.base+OFF_PTM)[8] * (T.base+OFF_PMA+100*(8))[0] * st.meso_tn1[4] * st.meso_tn1[4] / (pow(((T.base+OFF_PTM)[4] * (T.base+OFF_PTL+100*(3))[0]), 2.0));
    }
    const double g28 = fl.sw[21] * G(st, 4, (T.base+OFF_PD+150*(2)), in, fl);
    const double zhf = (T.base+OFF_PDL+25*(1))[24] * (1.0 + fl.sw[5] * (T.base+OFF_PDL+25*(0))[24] * sin(dgtr * in.g_lat) * cos(dr * (in.doy - (T.base+OFF_PT)[13])));
    out.t[0] = tinf;
    const double xmm = (T.base+OFF_PDM+10*(2))[4];
    const double z = in.alt;
    double tz;
    double* tn1 = st.meso_tn1;
    double* tgn1 = st.meso_tgn1;

    // N2
    const double db28 = (T.base+OFF_PDM+10*(2))[0] * exp(g28) * (T.base+OFF_PD+150*(2))[0];
    out.d[2] = densu(st, z, db28, tinf, tlb, 28.0, alpha[2], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    const double zh28 = (T.base+OFF_PDM+10*(2))[2] * zhf;
    const double zhm28 = (T.base+OFF_PDM+10*(2))[3] * (T.base+OFF_PDL+25*(1))[5];
    const double xmd = 28.0 - xmm;
    double b28 = 0.0;
    if (!(z > altl[7] && z >= zn1[0]))
        b28 = densu(st, zh28, db28, tinf, tlb, xmd, (alpha[2] - 1.0), &tz, (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[2])) {
        st.dm28 = densu(st, z, b28, tinf, tlb, xmm, alpha[2], &tz, (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[2] = dnet(out.d[2], st.dm28, zhm28, xmm, 28.0);
    }

    // He
    const double g4 = fl.sw[21] * G(st, 2, (T.base+OFF_PD+150*(0)), in, fl);
    const double db04 = (T.base+OFF_PDM+10*(0))[0] * exp(g4) * (T.base+OFF_PD+150*(0))[0];
    out.d[0] = densu(st, z, db04, tinf, tlb, 4., alpha[0], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z < altl[0])) {
        const double zh04 = (T.base+OFF_PDM+10*(0))[2];
        const double b04 = densu(st, zh04, db04, tinf, tlb, 4. - xmm, alpha[0] - 1., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        st.dm04 = densu(st, z, b04, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[0] = dnet(out.d[0], st.dm04, zhm28, xmm, 4.);
        const double rl = log(b28 * (T.base+OFF_PDM+10*(0))[1] / b04);
        const double zc04 = (T.base+OFF_PDM+10*(0))[4] * (T.base+OFF_PDL+25*(1))[0];
        const double hc04 = (T.base+OFF_PDM+10*(0))[5] * (T.base+OFF_PDL+25*(1))[1];
        out.d[0] = out.d[0] * ccor(z, rl, hc04, zc04);
    }

    // O
    const double g16 = fl.sw[21] * G(st, 3, (T.base+OFF_PD+150*(1)), in, fl);
    const double db16 = (T.base+OFF_PDM+10*(1))[0] * exp(g16) * (T.base+OFF_PD+150*(1))[0];
    out.d[1] = densu(st, z, db16, tinf, tlb, 16., alpha[1], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[1])) {
        const double zh16 = (T.base+OFF_PDM+10*(1))[2];
        const double b16 = densu(st, zh16, db16, tinf, tlb, 16.0 - xmm, (alpha[1] - 1.0), &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        st.dm16 = densu(st, z, b16, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[1] = dnet(out.d[1], st.dm16, zhm28, xmm, 16.);
        const double rl = (T.base+OFF_PDM+10*(1))[1] * (T.base+OFF_PDL+25*(1))[16] * (1.0 + fl.sw[1] * (T.base+OFF_PDL+25*(0))[23] * (in.f107A - 150.0));
        const double hc16 = (T.base+OFF_PDM+10*(1))[5] * (T.base+OFF_PDL+25*(1))[3];
        const double zc16 = (T.base+OFF_PDM+10*(1))[4] * (T.base+OFF_PDL+25*(1))[2];
        const double hc216 = (T.base+OFF_PDM+10*(1))[5] * (T.base+OFF_PDL+25*(1))[4];
        out.d[1] = out.d[1] * ccor2(z, rl, hc16, zc16, hc216);
        const double hcc16 = (T.base+OFF_PDM+10*(1))[7] * (T.base+OFF_PDL+25*(1))[13];
        const double zcc16 = (T.base+OFF_PDM+10*(1))[6] * (T.base+OFF_PDL+25*(1))[12];
        const double rc16 = (T.base+OFF_PDM+10*(1))[3] * (T.base+OFF_PDL+25*(1))[14];
        out.d[1] = out.d[1] * ccor(z, rc16, hcc16, zcc16);
    }

    // O2
    const double g32 = fl.sw[21] * G(st, 6, (T.base+OFF_PD+150*(4)), in, fl);
    const double db32 = (T.base+OFF_PDM+10*(3))[0] * exp(g32) * (T.base+OFF_PD+150*(4))[0];
    out.d[3] = densu(st, z, db32, tinf, tlb, 32., alpha[3], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if (fl.sw[15]) {
        if (z <= altl[3]) {
            const double zh32 = (T.base+OFF_PDM+10*(3))[2];
            const double b32 = densu(st, zh32, db32, tinf, tlb, 32. - xmm, alpha[3] - 1., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
            st.dm32 = densu(st, z, b32, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
            out.d[3] = dnet(out.d[3], st.dm32, zhm28, xmm, 32.);
            const double rl = log(b28 * (T.base+OFF_PDM+10*(3))[1] / b32);
            const double hc32 = (T.base+OFF_PDM+10*(3))[5] * (T.base+OFF_PDL+25*(1))[7];
            const double zc32 = (T.base+OFF_PDM+10*(3))[4] * (T.base+OFF_PDL+25*(1))[6];
            out.d[3] = out.d[3] * ccor(z, rl, hc32, zc32);
        }
        const double hcc32 = (T.base+OFF_PDM+10*(3))[7] * (T.base+OFF_PDL+25*(1))[22];
        const double hcc232 = (T.base+OFF_PDM+10*(3))[7] * (T.base+OFF_PDL+25*(0))[22];
        const double zcc32 = (T.base+OFF_PDM+10*(3))[6] * (T.base+OFF_PDL+25*(1))[21];
        const double rc32 = (T.base+OFF_PDM+10*(3))[3] * (T.base+OFF_PDL+25*(1))[23] * (1. + fl.sw[1] * (T.base+OFF_PDL+25*(0))[23] * (in.f107A - 150.));
        out.d[3] = out.d[3] * ccor2(z, rc32, hcc32, zcc32, hcc232);
    }

    // Ar
    const double g40 = fl.sw[21] * G(st, 7, (T.base+OFF_PD+150*(5)), in, fl);
    const double db40 = (T.base+OFF_PDM+10*(4))[0] * exp(g40) * (T.base+OFF_PD+150*(5))[0];
    out.d[4] = densu(st, z, db40, tinf, tlb, 40., alpha[4], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[4])) {
        const double zh40 = (T.base+OFF_PDM+10*(4))[2];
        const double b40 = densu(st, zh40, db40, tinf, tlb, 40. - xmm, alpha[4] - 1., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        st.dm40 = densu(st, z, b40, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[4] = dnet(out.d[4], st.dm40, zhm28, xmm, 40.);
        const double rl = log(b28 * (T.base+OFF_PDM+10*(4))[1] / b40);
        const double hc40 = (T.base+OFF_PDM+10*(4))[5] * (T.base+OFF_PDL+25*(1))[9];
        const double zc40 = (T.base+OFF_PDM+10*(4))[4] * (T.base+OFF_PDL+25*(1))[8];
        out.d[4] = out.d[4] * ccor(z, rl, hc40, zc40);
    }

    // H
    const double g1 = fl.sw[21] * G(st, 8, (T.base+OFF_PD+150*(6)), in, fl);
    const double db01 = (T.base+OFF_PDM+10*(5))[0] * exp(g1) * (T.base+OFF_PD+150*(6))[0];
    out.d[6] = densu(st, z, db01, tinf, tlb, 1., alpha[6], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[6])) {
        const double zh01 = (T.base+OFF_PDM+10*(5))[2];
        const double b01 = densu(st, zh01, db01, tinf, tlb, 1. - xmm, alpha[6] - 1., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        st.dm01 = densu(st, z, b01, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[6] = dnet(out.d[6], st.dm01, zhm28, xmm, 1.);
        const double rl = log(b28 * (T.base+OFF_PDM+10*(5))[1] * sqrt((T.base+OFF_PDL+25*(1))[17] * (T.base+OFF_PDL+25*(1))[17]) / b01);
        const double hc01 = (T.base+OFF_PDM+10*(5))[5] * (T.base+OFF_PDL+25*(1))[11];
        const double zc01 = (T.base+OFF_PDM+10*(5))[4] * (T.base+OFF_PDL+25*(1))[10];
        out.d[6] = out.d[6] * ccor(z, rl, hc01, zc01);
        const double hcc01 = (T.base+OFF_PDM+10*(5))[7] * (T.base+OFF_PDL+25*(1))[19];
        const double zcc01 = (T.base+OFF_PDM+10*(5))[6] * (T.base+OFF_PDL+25*(1))[18];
        const double rc01 = (T.base+OFF_PDM+10*(5))[3] * (T.base+OFF_PDL+25*(1))[20];
        out.d[6] = out.d[6] * ccor(z, rc01, hcc01, zcc01);
    }

    // N
    const double g14 = fl.sw[21] * G(st, 9, (T.base+OFF_PD+150*(7)), in, fl);
    const double db14 = (T.base+OFF_PDM+10*(6))[0] * exp(g14) * (T.base+OFF_PD+150*(7))[0];
    out.d[7] = densu(st, z, db14, tinf, tlb, 14., alpha[7], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if ((fl.sw[15]) && (z <= altl[7])) {
        const double zh14 = (T.base+OFF_PDM+10*(6))[2];
        const double b14 = densu(st, zh14, db14, tinf, tlb, 14. - xmm, alpha[7] - 1., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        st.dm14 = densu(st, z, b14, tinf, tlb, xmm, 0., &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
        out.d[7] = dnet(out.d[7], st.dm14, zhm28, xmm, 14.);
        const double rl = log(b28 * (T.base+OFF_PDM+10*(6))[1] * sqrt((T.base+OFF_PDL+25*(0))[2] * (T.base+OFF_PDL+25*(0))[2]) / b14);
        const double hc14 = (T.base+OFF_PDM+10*(6))[5] * (T.base+OFF_PDL+25*(0))[1];
        const double zc14 = (T.base+OFF_PDM+10*(6))[4] * (T.base+OFF_PDL+25*(0))[0];
        out.d[7] = out.d[7] * ccor(z, rl, hc14, zc14);
        const double hcc14 = (T.base+OFF_PDM+10*(6))[7] * (T.base+OFF_PDL+25*(0))[4];
        const double zcc14 = (T.base+OFF_PDM+10*(6))[6] * (T.base+OFF_PDL+25*(0))[3];
        const double rc14 = (T.base+OFF_PDM+10*(6))[3] * (T.base+OFF_PDL+25*(0))[5];
        out.d[7] = out.d[7] * ccor(z, rc14, hcc14, zcc14);
    }

    // anomalous O
    const double g16h = fl.sw[21] * G(st, 10, (T.base+OFF_PD+150*(8)), in, fl);
    const double db16h = (T.base+OFF_PDM+10*(7))[0] * exp(g16h) * (T.base+OFF_PD+150*(8))[0];
    const double tho = (T.base+OFF_PDM+10*(7))[9] * (T.base+OFF_PDL+25*(0))[6];
    const double ddo = densu(st, z, db16h, tho, tho, 16., alpha[8], &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    const double zsht = (T.base+OFF_PDM+10*(7))[5];
    const double zmho = (T.base+OFF_PDM+10*(7))[4];
    const double zsho = scalh(st, zmho, 16.0, tho);
    out.d[8] = ddo * exp(-zsht / zsho * (exp(-(z - zmho) / zsht) - 1.));

    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] + 14.0 * out.d[7]);
    const double zt = sqrt(in.alt * in.alt);
    (void)densu(st, zt, 1.0, tinf, tlb, 0.0, 0.0, &out.t[1], (T.base+OFF_PTM)[5], s, mn1, zn1, tn1, tgn1);
    if (fl.sw[0]) {
        for (int i = 0; i < 9; ++i) out.d[i] = out.d[i] * 1.0E6;
        out.d[5] = out.d[5] / 1000;
    }
}

// :958-1065

// ---------------------------------------------------------------------------
// gts7 split for the GPU: the common prefix (exospheric temperature, lower
// boundary temperature and gradient, lower thermosphere nodes) and one
// function per species, so the waves of a tile can evaluate the species
// concurrently. Same statements as gts7 above (Nrlmsise00Model.inl:1203-1498);
// each species recomputes the N2 mixed density at Zlb it consumes. Densu's
// writes to the lower thermosphere nodes depend only on (tinf, tlb, s), so the
// species are order-independent. Valid for alt >= zn2[0] = 72.5 km.
struct Pre {
    double tinf, tlb, s, zhf, xmm;
    double zn1[5];
};

template <class Globe>
__device__ inline void gts7Prefix(State& st, const NrlTables& T, Input& in, const Flags& fl, const Globe& G, Pre& pre) {
    const double* H = T.hot;
    double zn1[5] = {120.0, 110.0, 100.0, 90.0, 72.5};
    const double dgtr = 1.74533E-2, dr = 1.72142E-2;
    const double za = H[HOT_PDL + 25 + 15];
    zn1[0] = za;
    st.g0Valid = false;
    G.basis(st, in, fl);
    double tinf;
    if (in.alt > zn1[0]) tinf = H[HOT_PTM + 0] * H[HOT_PT + 0] * (1.0 + fl.sw[16] * G(st, 0, (T.base+OFF_PT), in, fl));
    else tinf = H[HOT_PTM + 0] * H[HOT_PT + 0];
    double g0v;
    if (in.alt > zn1[4]) g0v = H[HOT_PTM + 3] * H[HOT_PS0] * (1.0 + fl.sw[19] * G(st, 1, (T.base+OFF_PS), in, fl));
    else g0v = H[HOT_PTM + 3] * H[HOT_PS0];
    const double tlb = H[HOT_PTM + 1] * (1.0 + fl.sw[17] * G(st, 5, (T.base+OFF_PD+150*(3)), in, fl)) * H[HOT_PD0 + 3];
    const double s = g0v / (tinf - tlb);
    if (in.alt < 300.0) {
        st.meso_tn1[1] = H[HOT_PTM + 6] * H[HOT_PTL0 + 0] / (1.0 - fl.sw[18] * G.lower(st, 0, (T.base+OFF_PTL+100*(0)), in, fl));
        st.meso_tn1[2] = H[HOT_PTM + 2] * H[HOT_PTL0 + 1] / (1.0 - fl.sw[18] * G.lower(st, 1, (T.base+OFF_PTL+100*(1)), in, fl));
        st.meso_tn1[3] = H[HOT_PTM + 7] * H[HOT_PTL0 + 2] / (1.0 - fl.sw[18] * G.lower(st, 2, (T.base+OFF_PTL+100*(2)), in, fl));
        st.meso_tn1[4] = H[HOT_PTM + 4] * H[HOT_PTL0 + 3] / (1.0 - fl.sw[18] * fl.sw[20] * G.lower(st, 3, (T.base+OFF_PTL+100*(3)), in, fl));
        st.meso_tgn1[1] = H[HOT_PTM + 8] * H[HOT_PMA80] * (1.0 + fl.sw[18] * fl.sw[20] * G.lower(st, 4, (T.base+OFF_PMA+100*(8)), in, fl)) * st.meso_tn1[4] *
                          st.meso_tn1[4] / (pow((H[HOT_PTM + 4] * H[HOT_PTL0 + 3]), 2.0));
    } else {
        st.meso_tn1[1] = H[HOT_PTM + 6] * H[HOT_PTL0 + 0];
        st.meso_tn1[2] = H[HOT_PTM + 2] * H[HOT_PTL0 + 1];
        st.meso_tn1[3] = H[HOT_PTM + 7] * H[HOT_PTL0 + 2];
        st.meso_tn1[4] = H[HOT_PTM + 4] * H[HOT_PTL0 + 3];
        st.meso_tgn1[1] = H[HOT_PTM + 8] * H[HOT_PMA80] * st.meso_tn1[4] * st.meso_tn1[4] / (pow((H[HOT_PTM + 4] * H[HOT_PTL0 + 3]), 2.0));
    }
    pre.tinf = tinf;
    pre.tlb = tlb;
    pre.s = s;
    pre.zhf = H[HOT_PDL + 25 + 24] * (1.0 + fl.sw[5] * H[HOT_PDL + 24] * sin(dgtr * in.g_lat) * cos(dr * (in.doy - H[HOT_PT + 13])));
    pre.xmm = H[HOT_PDM + 24];
    for (int k = 0; k < 5; ++k) pre.zn1[k] = zn1[k];
}

// Per-species constants of gts7 (Nrlmsise00Model.inl:1203-1498) for species
// 0 N2, 1 He, 2 O, 3 O2, 4 Ar, 5 H, 6 N: pd block, globe7 table slot (FastGlobe
// epoch block), pdm row, molecular mass, thermal-diffusion alpha, mixing
// altitude altl, and flat pdl indices (pdl[0][i] = i, pdl[1][i] = 25 + i) of the
// chemistry/dissociation correction terms (-1: the species has none).
struct SpeciesPar {
    int pd, globe, row;
    double xm, alpha, altl;
    int fIdx, hIdx, zIdx, rcIdx, hccIdx, zccIdx;
};
static __constant__ SpeciesPar kSpeciesPar[7] = {
    {2, 4, 2, 28.0, 0.0, 160.0, -1, -1, -1, -1, -1, -1},   // N2
    {0, 2, 0, 4.0, -0.38, 200.0, -1, 26, 25, -1, -1, -1},  // He
    {1, 3, 1, 16.0, 0.0, 300.0, -1, 28, 27, 39, 38, 37},   // O  (ccor2: h1 pdl[1][3], zh pdl[1][2], h2 pdl[1][4])
    {4, 6, 3, 32.0, 0.0, 250.0, -1, 32, 31, -1, -1, -1},   // O2
    {5, 7, 4, 40.0, 0.17, 240.0, -1, 34, 33, -1, -1, -1},  // Ar
    {6, 8, 5, 1.0, -0.38, 320.0, 42, 36, 35, 45, 44, 43},  // H
    {7, 9, 6, 14.0, 0.0, 450.0, 2, 1, 0, 5, 4, 3},         // N
};

// Quantities that every densu call of a lane's species shares (same tinf, tlb,
// s and lower-boundary nodes), each computed once with densu's own statements
// (Nrlmsise00Model.inl:476-576), so a cached call returns the same bits as a
// full one:
//  * at the lane altitude z >= za: zeta(z, zlb) and the Bates temperature;
//  * below za (z clamped to za): the temperature at za, the lower-thermosphere
//    node update (tn1[0], tgn1[0]) and the five-node spline of 1/T;
//  * the N2 diffusive-equilibrium factor of the mixing densities at z
//    (xm = xmm, alpha = 0), shared by every species' mixed density;
//  * the N2 density at its turbopause height (b28) and at Zlb (db28).
struct DensuShared {
    double re, za, zlb, s2, tinf, tlb, glb, xmm;
    double zLane, zg2L, ttL, lnL;   // lnL = log(tlb / ttL)
    bool laneUpper;
    double zg2A, ttA, lnA, z1, z2, t1, zgdif, glb1;
    double xs[5], ys[5], y2out[5];
    double pmm, emm;   // pow(tlb/tt, 1 + gamma(xmm)) and its exp factor at the lane altitude
    double db28, b28;
};

// zeta(zz, z1) with the lane's Earth radius
__device__ inline double zeta0(const DensuShared& h, double zz) { return ((zz - h.z1) * (h.re + h.z1) / (h.re + zz)); }

// densu's diffusive-equilibrium density above za. The power (tlb/tt)^(1+alpha+gamma)
// is taken as exp((1+alpha+gamma) log(tlb/tt)) with the lane's shared logarithm
// (relative difference from pow ~1e-15).
__device__ inline double densuUpperFrom(const DensuShared& h, double zg2, double tt, double lnT, double dlb, double xm,
                                        double alpha) {
    const double rgas = 831.4;
    const double gamma = xm * h.glb / (h.s2 * rgas * h.tinf);
    double expl = exp(-h.s2 * gamma * zg2);
    if (expl > 50.0) expl = 50.0;
    if (tt <= 0) expl = 50.0;
    return dlb * exp(((1.0 + alpha + gamma)) * lnT) * expl;
}

// densuShare in three parts that depend on each other only through the common
// fields, so two waves can form the lane part and the lower-thermosphere part
// at once; together they perform densu's statements once each.
__device__ inline void densuShareCommon(DensuShared& h, const State& st, double alt, double tinf, double tlb, double s2,
                                        double zlb, double xmm, const double* zn1) {
    h.re = st.re;
    h.za = zn1[0];
    h.zlb = zlb;
    h.s2 = s2;
    h.tinf = tinf;
    h.tlb = tlb;
    h.xmm = xmm;
    h.glb = st.gsurf / pow((1.0 + zlb / st.re), 2.0);
    h.zLane = alt;
    h.laneUpper = !(alt < h.za);
}

// The lane altitude: zeta, the Bates temperature and its logarithm, and the
// mixing-density factor (xm = xmm, alpha = 0).
__device__ inline void densuShareLane(DensuShared& h, const State& st) {
    const double rgas = 831.4;
    const double zL = h.zLane > h.za ? h.zLane : h.za;
    h.zg2L = zeta(st, zL, h.zlb);
    h.ttL = h.tinf - (h.tinf - h.tlb) * exp(-h.s2 * h.zg2L);
    h.lnL = log(h.tlb / h.ttL);
    const double gamma = h.xmm * h.glb / (h.s2 * rgas * h.tinf);
    double expl = exp(-h.s2 * gamma * h.zg2L);
    if (expl > 50.0) expl = 50.0;
    if (h.ttL <= 0) expl = 50.0;
    h.pmm = exp(((1.0 + 0. + gamma)) * h.lnL);
    h.emm = expl;
}

// Below za: the temperature at za, the lower-thermosphere node update (tn1[0],
// tgn1[0]) and the five-node spline of 1/T.
__device__ inline void densuShareLower(DensuShared& h, const State& st, const double* zn1, double* tn1, double* tgn1) {
    constexpr int MN = 5;
    h.zg2A = zeta(st, h.za, h.zlb);
    h.ttA = h.tinf - (h.tinf - h.tlb) * exp(-h.s2 * h.zg2A);
    h.lnA = log(h.tlb / h.ttA);
    const double dta = (h.tinf - h.ttA) * h.s2 * pow(((st.re + h.zlb) / (st.re + h.za)), 2.0);
    tgn1[0] = dta;
    tn1[0] = h.ttA;
    h.z1 = zn1[0];
    h.z2 = zn1[MN - 1];
    h.t1 = tn1[0];
    const double t2 = tn1[MN - 1];
    h.zgdif = zeta(st, h.z2, h.z1);
#pragma unroll
    for (int k = 0; k < MN; ++k) {
        h.xs[k] = zeta(st, zn1[k], h.z1) / h.zgdif;
        h.ys[k] = 1.0 / tn1[k];
    }
    const double yd1 = -tgn1[0] / (h.t1 * h.t1) * h.zgdif;
    const double yd2 = -tgn1[1] / (t2 * t2) * h.zgdif * pow(((st.re + h.z2) / (st.re + h.z1)), 2.0);
    splineN(h.xs, h.ys, yd1, yd2, h.y2out);
    h.glb1 = st.gsurf / pow((1.0 + h.z1 / st.re), 2.0);
}

__device__ inline void densuShare(DensuShared& h, State& st, double alt, double tinf, double tlb, double s2, double zlb,
                                  double xmm, const double* zn1, double* tn1, double* tgn1) {
    densuShareCommon(h, st, alt, tinf, tlb, s2, zlb, xmm, zn1);
    densuShareLane(h, st);
    densuShareLower(h, st, zn1, tn1, tgn1);
    h.db28 = h.b28 = 0.0;
}

// densu at an altitude below za from the shared lower-thermosphere profile.
__device__ inline double densuBelow(const DensuShared& h, double alt, double dlb, double xm, double alpha) {
    const double rgas = 831.4;
    const double z = alt > h.z2 ? alt : h.z2;
    const double x = zeta0(h, z) / h.zgdif;
    double y;
    splintN(h.xs, h.ys, h.y2out, x, &y);
    const double tz = 1.0 / y;
    if (xm == 0) return tz;
    const double densa = densuUpperFrom(h, h.zg2A, h.ttA, h.lnA, dlb, xm, alpha);
    const double gamm = xm * h.glb1 * h.zgdif / rgas;
    double yi;
    spliniN(h.xs, h.ys, h.y2out, x, &yi);
    double expl = gamm * yi;
    if (expl > 50.0) expl = 50.0;
    if (tz <= 0) expl = 50.0;
    return densa * exp((1.0 + alpha) * log(h.t1 / tz)) * exp(-expl);
}

// densu at the lane altitude, or at another altitude through the cached
// profiles where they apply (full densu otherwise).
__device__ inline double densuLane(const DensuShared& h, State& st, double dlb, double xm, double alpha, const double* zn1,
                                   double* tn1, double* tgn1) {
    double tz;
    if (h.laneUpper) return densuUpperFrom(h, h.zg2L, h.ttL, h.lnL, dlb, xm, alpha);
    return densu(st, h.zLane, dlb, h.tinf, h.tlb, xm, alpha, &tz, h.zlb, h.s2, 5, zn1, tn1, tgn1);
}

__device__ inline double densuAt(const DensuShared& h, State& st, double alt, double dlb, double xm, double alpha,
                                 const double* zn1, double* tn1, double* tgn1) {
    double tz;
    if (alt < h.za) return densuBelow(h, alt, dlb, xm, alpha);
    return densu(st, alt, dlb, h.tinf, h.tlb, xm, alpha, &tz, h.zlb, h.s2, 5, zn1, tn1, tgn1);
}

// Species k (0 N2, 1 He, 2 O, 3 O2, 4 Ar, 5 H, 6 N, 7 anomalous O): number
// density in the model's cgs units (before the metric factor). One code path
// for species 0-6 driven by kSpeciesPar (k is uniform across the wave, so the
// parameters are scalar loads); each species performs the same operations in
// the same order as its block of gts7, with the densu calls served from the
// lane's DensuShared (gts7SpeciesShare). Densu's side effects
// (lower-thermosphere nodes, tz) depend only on (tinf, tlb, s), so the shared
// quantities may be evaluated before any species.
// The lane part of the shared quantities alone (densuShareLane's fields and the
// common ones), for a wave beside the one that runs gts7SpeciesShare<false>.
__device__ inline void gts7ShareLane(DensuShared& h, const State& st, const NrlTables& T, const Input& in, const Pre& pre) {
    densuShareCommon(h, st, in.alt, pre.tinf, pre.tlb, pre.s, T.hot[HOT_PTM + 5], pre.xmm, pre.zn1);
    densuShareLane(h, st);
}

// The N2 density at Zlb (db28) from its globe7 table, where a species needs it.
template <class Globe>
__device__ __attribute__((always_inline)) inline void gts7ShareDb28(DensuShared& h, bool needDb28, State& st, const NrlTables& T,
                                                                    Input& in, const Flags& fl, const Globe& G, const Pre& pre) {
    const double* pdm = T.hot + HOT_PDM;
    const double z = in.alt;
    const bool wantB28 = !(z > 450.0 && z >= pre.zn1[0]);
    if (wantB28 || needDb28) {
        const double g28 = fl.sw[21] * G(st, 4, (T.base + OFF_PD + 150 * (2)), in, fl);
        h.db28 = pdm[20] * exp(g28) * T.hot[HOT_PD0 + 2];
    }
}

// The N2 mixed density at its turbopause height (b28) from the shared
// lower-thermosphere quantities and db28; the last statement of
// gts7SpeciesShare, which B28 = false leaves out.
__device__ inline void gts7ShareB28(DensuShared& h, State& st, const NrlTables& T, const Input& in, const Pre& pre) {
    double zn1[5] = {pre.zn1[0], pre.zn1[1], pre.zn1[2], pre.zn1[3], pre.zn1[4]};
    const double* pdm = T.hot + HOT_PDM;
    const double z = in.alt;
    const bool wantB28 = !(z > 450.0 && z >= zn1[0]);
    if (wantB28) {
        const double zh28 = pdm[22] * pre.zhf;
        h.b28 = densuAt(h, st, zh28, h.db28, 28.0 - pre.xmm, (0.0 - 1.0), zn1, st.meso_tn1, st.meso_tgn1);
    }
}

// LANE = false leaves densuShareLane's fields to another wave (gts7ShareLane);
// B28 = false leaves b28 to gts7ShareB28 and DB28 = false also db28 (to
// gts7ShareDb28).
template <class Globe, bool LANE = true, bool B28 = true, bool DB28 = true>
__device__ __attribute__((always_inline)) inline void gts7SpeciesShare(DensuShared& h, bool needDb28, State& st,
                                                                       const NrlTables& T, Input& in, const Flags& fl,
                                                                       const Globe& G, const Pre& pre) {
    double zn1[5] = {pre.zn1[0], pre.zn1[1], pre.zn1[2], pre.zn1[3], pre.zn1[4]};
    const double z = in.alt;
    if constexpr (LANE) {
        densuShare(h, st, z, pre.tinf, pre.tlb, pre.s, T.hot[HOT_PTM + 5], pre.xmm, zn1, st.meso_tn1, st.meso_tgn1);
    } else {
        densuShareCommon(h, st, z, pre.tinf, pre.tlb, pre.s, T.hot[HOT_PTM + 5], pre.xmm, zn1);
        densuShareLower(h, st, zn1, st.meso_tn1, st.meso_tgn1);
        h.db28 = h.b28 = 0.0;
    }
    if constexpr (DB28) gts7ShareDb28(h, needDb28, st, T, in, fl, G, pre);
    if constexpr (B28) gts7ShareB28(h, st, T, in, pre);
}

// Species k of 0-6 up to the statements that need b28 (the N2 mixed density at
// its turbopause height): the diffusive and mixed densities, and every
// correction factor that does not depend on b28, either applied (when gts7
// applies it before the b28 factor, or when there is none) or kept (c1, c2) for
// the tail, which applies the rest in gts7's order. A factor of 1.0 leaves the
// density's bits unchanged.
struct SpeciesHead {
    double d;    // density so far
    double bk;   // the species' density at its turbopause height (tail 2)
    double c1, c2;   // factors after the b28 one (1.0: none)
    CcorPre cA;  // the altitude part of the b28 factor
    int tail;    // 0: d is final; 1: N2 mixing (dm28 and dnet); 2: ccor of log(b28 ... / bk), then c1, c2
};

// The correction factors of species k that depend on the lane altitude (and
// F10.7A) alone, so a wave can form them before the shared densu quantities
// exist: O's chemistry ccor2, the ccor of O, H and N with rcIdx, O2's ccor2 and
// the altitude part of the ccor that the tail applies (1.0 or unused when the
// species has none at this altitude).
struct SpeciesPre {
    double c2, rc, c32;
    CcorPre cA;
};

__device__ inline SpeciesPre gts7SpeciesPre(int k, const NrlTables& T, const Input& in, const Flags& fl) {
    SpeciesPre q{1.0, 1.0, 1.0, {0, 0.0}};
    const double z = in.alt;
    const double* pdm = T.hot + HOT_PDM;
    const double* pdl = T.hot + HOT_PDL;
    const SpeciesPar P = kSpeciesPar[k];
    const double* pdmk = pdm + 10 * P.row;
    const bool mix = fl.sw[15] && (k == 1 ? z < P.altl : z <= P.altl);
    if (mix && k != 0) {
        if (k == 2) {
            const double rl = pdmk[1] * pdl[25 + 16] * (1.0 + fl.sw[1] * pdl[23] * (in.f107A - 150.0));
            q.c2 = ccor2(z, rl, pdmk[5] * pdl[P.hIdx], pdmk[4] * pdl[P.zIdx], pdmk[5] * pdl[25 + 4]);
        } else {
            q.cA = ccorPre(z, pdmk[5] * pdl[P.hIdx], pdmk[4] * pdl[P.zIdx]);
        }
        if (P.rcIdx >= 0) q.rc = ccor(z, pdmk[3] * pdl[P.rcIdx], pdmk[7] * pdl[P.hccIdx], pdmk[6] * pdl[P.zccIdx]);
    }
    if (k == 3 && fl.sw[15]) {
        const double hcc32 = pdmk[7] * pdl[25 + 22];
        const double hcc232 = pdmk[7] * pdl[22];
        const double zcc32 = pdmk[6] * pdl[25 + 21];
        const double rc32 = pdmk[3] * pdl[25 + 23] * (1. + fl.sw[1] * pdl[23] * (in.f107A - 150.));
        q.c32 = ccor2(z, rc32, hcc32, zcc32, hcc232);
    }
    return q;
}

template <class Globe>
__device__ __attribute__((always_inline)) inline SpeciesHead gts7SpeciesHead(int k, const DensuShared& h, State& st,
                                                                             const NrlTables& T, Input& in, const Flags& fl,
                                                                             const Globe& G, const Pre& pre,
                                                                             const SpeciesPre& q) {
    const double xmm = pre.xmm;
    double zn1[5] = {pre.zn1[0], pre.zn1[1], pre.zn1[2], pre.zn1[3], pre.zn1[4]};
    const double z = in.alt;
    double* tn1 = st.meso_tn1;
    double* tgn1 = st.meso_tgn1;
    const double* pdm = T.hot + HOT_PDM;   // pdm[row][i] = pdm[10 row + i]
    const double* pdl = T.hot + HOT_PDL;   // pdl[blk][i] = pdl[25 blk + i]
    SpeciesHead r{0.0, 0.0, 1.0, 1.0, {0, 0.0}, 0};
    const SpeciesPar P = kSpeciesPar[k];
    const double* pdkG = T.base + OFF_PD + 150 * P.pd;
    const double* pdmk = pdm + 10 * P.row;
    double dbk = h.db28;
    if (k != 0) {
        const double gk = fl.sw[21] * G(st, P.globe, pdkG, in, fl);
        dbk = pdmk[0] * exp(gk) * T.hot[HOT_PD0 + P.pd];
    }
    double d = densuLane(h, st, dbk, P.xm, P.alpha, zn1, tn1, tgn1);
    const double zhm28 = pdm[23] * pdl[25 + 5];
    const bool mix = fl.sw[15] && (k == 1 ? z < P.altl : z <= P.altl);
    if (mix) {
        // mixed density at the lane altitude: densu(z, b, ..., xmm, 0)
        if (k == 0) {
            r.tail = 1;
        } else {
            const double zh = pdmk[2];
            const double bk = densuAt(h, st, zh, dbk, P.xm - xmm, P.alpha - 1.0, zn1, tn1, tgn1);
            const double dmk = h.laneUpper ? bk * h.pmm * h.emm : densuLane(h, st, bk, xmm, 0., zn1, tn1, tgn1);
            d = dnetFast(d, dmk, zhm28, xmm, P.xm);
            if (k == 2) {
                d = d * q.c2;
            } else {
                r.bk = bk;
                r.cA = q.cA;
                r.tail = 2;
            }
            if (P.rcIdx >= 0) {
                if (r.tail == 2) r.c1 = q.rc;
                else d = d * q.rc;
            }
        }
    }
    if (k == 3 && fl.sw[15]) {
        if (r.tail == 2) r.c2 = q.c32;
        else d = d * q.c32;
    }
    r.d = d;
    return r;
}

// The rest of species k from its head and b28 (h.b28).
__device__ __attribute__((always_inline)) inline double gts7SpeciesTail(int k, const SpeciesHead& r, const DensuShared& h,
                                                                       State& st, const NrlTables& T, const Pre& pre) {
    if (r.tail == 0) return r.d;
    const double xmm = pre.xmm;
    const double* pdm = T.hot + HOT_PDM;
    const double* pdl = T.hot + HOT_PDL;
    const double zhm28 = pdm[23] * pdl[25 + 5];
    const double b28 = h.b28;
    if (r.tail == 1) {
        double zn1[5] = {pre.zn1[0], pre.zn1[1], pre.zn1[2], pre.zn1[3], pre.zn1[4]};
        const double dm28 = h.laneUpper ? b28 * h.pmm * h.emm : densuLane(h, st, b28, xmm, 0.0, zn1, st.meso_tn1, st.meso_tgn1);
        return dnetFast(r.d, dm28, zhm28, xmm, 28.0);
    }
    const SpeciesPar P = kSpeciesPar[k];
    const double* pdmk = pdm + 10 * P.row;
    const double f = P.fIdx >= 0 ? sqrt(pdl[P.fIdx] * pdl[P.fIdx]) : 1.0;
    const double rl = P.fIdx >= 0 ? log(b28 * pdmk[1] * f / r.bk) : log(b28 * pdmk[1] / r.bk);
    double d = r.d * ccorFrom(r.cA, rl);
    d = d * r.c1;
    d = d * r.c2;
    return d;
}

template <class Globe>
__device__ __attribute__((always_inline)) inline double gts7Species(int k, const DensuShared& h, State& st, const NrlTables& T,
                                                                    Input& in, const Flags& fl, const Globe& G, const Pre& pre) {
    if (k == 7) {  // anomalous O (d[8])
        const int mn1 = 5;
        const double s = pre.s;
        double zn1[5] = {pre.zn1[0], pre.zn1[1], pre.zn1[2], pre.zn1[3], pre.zn1[4]};
        const double z = in.alt;
        double tz;
        const double* pdm = T.hot + HOT_PDM;
        const double* pdl = T.hot + HOT_PDL;
        const double zlbT = T.hot[HOT_PTM + 5];
        const double g16h = fl.sw[21] * G(st, 10, (T.base + OFF_PD + 150 * (8)), in, fl);
        const double db16h = pdm[70] * exp(g16h) * T.hot[HOT_PD0 + 8];
        const double tho = pdm[79] * pdl[6];
        const double ddo = densu(st, z, db16h, tho, tho, 16., 0.0, &tz, zlbT, s, mn1, zn1, st.meso_tn1, st.meso_tgn1);
        const double zsht = pdm[75];
        const double zmho = pdm[74];
        const double zsho = scalh(st, zmho, 16.0, tho);
        return ddo * exp(-zsht / zsho * (exp(-(z - zmho) / zsht) - 1.));
    }
    const SpeciesPre q = gts7SpeciesPre(k, T, in, fl);
    const SpeciesHead r = gts7SpeciesHead(k, h, st, T, in, fl, G, pre, q);
    return gts7SpeciesTail(k, r, h, st, T, pre);
}

// Mass-density weights of the species in gtd7d (amu): d[5] = 1.66e-24 sum w_k d_k.
__device__ inline double speciesMass(int k) {
    const double w[8] = {28.0, 4.0, 16.0, 32.0, 40.0, 1.0, 14.0, 16.0};
    return w[k];
}

template <class Globe>
__device__ inline void gtd7(State& st, const NrlTables& T, Input& in, Flags& fl, Output& out, const Globe& G) {
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
    const double xmm = (T.base+OFF_PDM+10*(2))[4];
    double altt = in.alt > zn2[0] ? in.alt : zn2[0];
    const double tmp = in.alt;
    in.alt = altt;
    gts7(st, T, in, fl, so, G);
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
    st.meso_tn2[1] = (T.base+OFF_PMA+100*(0))[0] * (T.base+OFF_PAVGM)[0] / (1.0 - fl.sw[20] * glob7s(st, (T.base+OFF_PMA+100*(0)), in, fl));
    st.meso_tn2[2] = (T.base+OFF_PMA+100*(1))[0] * (T.base+OFF_PAVGM)[1] / (1.0 - fl.sw[20] * glob7s(st, (T.base+OFF_PMA+100*(1)), in, fl));
    st.meso_tn2[3] = (T.base+OFF_PMA+100*(2))[0] * (T.base+OFF_PAVGM)[2] / (1.0 - fl.sw[20] * fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(2)), in, fl));
    st.meso_tgn2[1] = (T.base+OFF_PAVGM)[8] * (T.base+OFF_PMA+100*(9))[0] * (1.0 + fl.sw[20] * fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(9)), in, fl)) * st.meso_tn2[3] *
                      st.meso_tn2[3] / (pow(((T.base+OFF_PMA+100*(2))[0] * (T.base+OFF_PAVGM)[2]), 2.0));
    st.meso_tn3[0] = st.meso_tn2[3];
    if (in.alt <= zn3[0]) {
        st.meso_tgn3[0] = st.meso_tgn2[1];
        st.meso_tn3[1] = (T.base+OFF_PMA+100*(3))[0] * (T.base+OFF_PAVGM)[3] / (1.0 - fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(3)), in, fl));
        st.meso_tn3[2] = (T.base+OFF_PMA+100*(4))[0] * (T.base+OFF_PAVGM)[4] / (1.0 - fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(4)), in, fl));
        st.meso_tn3[3] = (T.base+OFF_PMA+100*(5))[0] * (T.base+OFF_PAVGM)[5] / (1.0 - fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(5)), in, fl));
        st.meso_tn3[4] = (T.base+OFF_PMA+100*(6))[0] * (T.base+OFF_PAVGM)[6] / (1.0 - fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(6)), in, fl));
        st.meso_tgn3[1] = (T.base+OFF_PMA+100*(7))[0] * (T.base+OFF_PAVGM)[7] * (1.0 + fl.sw[22] * glob7s(st, (T.base+OFF_PMA+100*(7)), in, fl)) * st.meso_tn3[4] *
                          st.meso_tn3[4] / (pow(((T.base+OFF_PMA+100*(6))[0] * (T.base+OFF_PAVGM)[6]), 2.0));
    }
    double dmc = 0;
    if (in.alt > zmix) dmc = 1.0 - (zn2[0] - in.alt) / (zn2[0] - zmix);
    const double dz28 = so.d[2];
    double tz;
    double dmr = so.d[2] / dm28m - 1.0;
    out.d[2] = densm(st, in.alt, dm28m, xmm, &tz, mn3, zn3, st.meso_tn3, st.meso_tgn3, mn2, zn2, st.meso_tn2, st.meso_tgn2);
    out.d[2] = out.d[2] * (1.0 + dmr * dmc);
    dmr = so.d[0] / (dz28 * (T.base+OFF_PDM+10*(0))[1]) - 1.0;
    out.d[0] = out.d[2] * (T.base+OFF_PDM+10*(0))[1] * (1.0 + dmr * dmc);
    out.d[1] = 0;
    out.d[8] = 0;
    dmr = so.d[3] / (dz28 * (T.base+OFF_PDM+10*(3))[1]) - 1.0;
    out.d[3] = out.d[2] * (T.base+OFF_PDM+10*(3))[1] * (1.0 + dmr * dmc);
    dmr = so.d[4] / (dz28 * (T.base+OFF_PDM+10*(4))[1]) - 1.0;
    out.d[4] = out.d[2] * (T.base+OFF_PDM+10*(4))[1] * (1.0 + dmr * dmc);
    out.d[6] = 0;
    out.d[7] = 0;
    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] + 14.0 * out.d[7]);
    if (fl.sw[0]) out.d[5] = out.d[5] / 1000;
    st.dd = densm(st, in.alt, 1.0, 0, &tz, mn3, zn3, st.meso_tn3, st.meso_tgn3, mn2, zn2, st.meso_tn2, st.meso_tgn2);
    out.t[1] = tz;
}

// :1073-1078
template <class Globe>
__device__ inline void gtd7d(State& st, const NrlTables& T, Input& in, Flags& fl, Output& out, const Globe& G) {
    gtd7(st, T, in, fl, out, G);
    out.d[5] = 1.66E-24 * (4.0 * out.d[0] + 16.0 * out.d[1] + 28.0 * out.d[2] + 32.0 * out.d[3] + 40.0 * out.d[4] + out.d[6] +
                           14.0 * out.d[7] + 16.0 * out.d[8]);
    if (fl.sw[0]) out.d[5] = out.d[5] / 1000;
}


// ---------------------------------------------------------------------------
// Policies for the globe7 evaluations inside gts7.
// GenericGlobe: the model as published (any switches) — used by the test probe.
struct GenericGlobe {
    __device__ void basis(State& st, const Input& in, const Flags& fl) const { prepareGlobeBasis(st, in, fl); }
    __device__ double operator()(State& st, int, const double* p, const Input& in, const Flags& fl) const {
        return globe7(st, p, in, fl);
    }
    // glob7s of lower-thermosphere table j (lowerTable order)
    __device__ double lower(State& st, int, const double* p, const Input& in, const Flags& fl) const {
        return glob7s(st, p, in, fl);
    }
};

// Epoch-only parts of globe7 for one coefficient table with the propagator's
// switch setting (all on, switch 9 = -1): day-of-year cosines, flux factors,
// UT cosines, the 3-hour Ap function where it does not depend on latitude
// (p[138] = 0), and the g0 values where it does. 18 values per table.
enum : int {
    GE_CD14 = 0, GE_CD18, GE_CD32, GE_CD39, GE_F1, GE_F2, GE_T0, GE_UT71, GE_UT79C, GE_UT79S, GE_UT58, GE_APT0,
    GE_G0 = 12,  // 6 values
    GE_STRIDE = 18,
    GE_TABLES = 11,
    GE_DFA = GE_TABLES * GE_STRIDE,
    GE_COUNT = GE_DFA + 1
};

// Table j: 0 = pt, 1 = ps, 2 + k = pd[k].
__device__ inline const double* globeTable(const NrlTables& T, int j) {
    return j == 0 ? T.base + OFF_PT : (j == 1 ? T.base + OFF_PS : T.base + OFF_PD + 150 * (j - 2));
}

// Fills out[GE_COUNT] from the epoch inputs (doy, sec, f107, f107A, ap[7]).
__device__ inline void globeEpoch(const NrlTables& T, int doy, double sec, double f107, double f107A, const double* ap,
                                  double* out) {
    const double sr = 7.2722E-5, dr = 1.72142E-2;
    const double df = f107 - f107A;
    const double dfa = f107A - 150.0;
    out[GE_DFA] = dfa;
    for (int j = 0; j < GE_TABLES; ++j) {
        const double* p = globeTable(T, j);
        double* e = out + GE_STRIDE * j;
        e[GE_CD32] = cos(dr * (doy - p[31]));
        e[GE_CD18] = cos(2.0 * dr * (doy - p[17]));
        e[GE_CD14] = cos(dr * (doy - p[13]));
        e[GE_CD39] = cos(2.0 * dr * (doy - p[38]));
        e[GE_T0] = p[19] * df * (1.0 + p[59] * dfa) + p[20] * df * df + p[21] * dfa + p[29] * pow(dfa, 2.0);
        e[GE_F1] = 1.0 + (p[47] * dfa + p[19] * df + p[20] * df * df);
        e[GE_F2] = 1.0 + (p[49] * dfa + p[19] * df + p[20] * df * df);
        e[GE_UT71] = cos(sr * (sec - p[71]));
        e[GE_UT79C] = cos(sr * (sec - p[79]));
        e[GE_UT79S] = sin(sr * (sec - p[79]));
        e[GE_UT58] = cos(sr * (sec - p[58]));
        for (int k = 0; k < 6; ++k) e[GE_G0 + k] = g0(ap[k + 1], p);
        e[GE_APT0] = 0.0;
        if (p[51] != 0 && p[138] == 0.0) {
            double exp1 = exp(-10800.0 * sqrt(p[51] * p[51]) / 1.0);
            if (exp1 > 0.99999) exp1 = 0.99999;
            const double* g = e + GE_G0;
            e[GE_APT0] = (g[0] + (g[1] * exp1 + g[2] * exp1 * exp1 + g[3] * pow(exp1, 3.0) +
                                  (g[4] * pow(exp1, 4.0) + g[5] * pow(exp1, 12.0)) * (1.0 - pow(exp1, 8.0)) / (1.0 - exp1))) /
                         sumex(exp1);
        }
    }
}

// Lower-thermosphere node tables of gts7 (glob7s, Nrlmsise00Model.inl:858-950):
// ptl[0..3] and pma[8]. Their day-of-year cosines and the day-of-year factors of
// the longitudinal term are epoch-only: per table cd14, cd18, cd32, cd39 and
// LX = p[80] cos(dr (doy - p[81])) + p[85] cos(2 dr (doy - p[86])),
// LY = p[83] cos(dr (doy - p[84])) + p[87] cos(2 dr (doy - p[88]))
// with the propagator's switch setting (swc = 1).
enum : int { GL_CD14 = 0, GL_CD18, GL_CD32, GL_CD39, GL_LX, GL_LY, GL_STRIDE = 6, GL_TABLES = 5, GL_COUNT = GL_TABLES * GL_STRIDE };
static_assert(GE_COUNT == 199 && GL_COUNT == 30, "epoch record NRLMSISE-00 block layout");

__device__ inline const double* lowerTable(const NrlTables& T, int j) {
    return T.base + (j < 4 ? OFF_PTL + 100 * j : OFF_PMA + 100 * 8);
}

__device__ inline void lowerEpoch(const NrlTables& T, int doy, double* out) {
    const double dr = 1.72142E-2;
    for (int j = 0; j < GL_TABLES; ++j) {
        const double* p = lowerTable(T, j);
        double* e = out + GL_STRIDE * j;
        e[GL_CD32] = cos(dr * (doy - p[31]));
        e[GL_CD18] = cos(2.0 * dr * (doy - p[17]));
        e[GL_CD14] = cos(dr * (doy - p[13]));
        e[GL_CD39] = cos(2.0 * dr * (doy - p[38]));
        e[GL_LX] = p[80] * cos(dr * (doy - p[81])) + p[85] * cos(2.0 * dr * (doy - p[86]));
        e[GL_LY] = p[83] * cos(dr * (doy - p[84])) + p[87] * cos(2.0 * dr * (doy - p[88]));
    }
}

// glob7s of table p with its epoch terms e (GL_STRIDE values) and the lane basis
// of State, for the propagator's switch setting: the published terms in the
// published summation order (t[9], t[11], t[12] are zero).
__device__ __attribute__((always_inline)) inline double glob7sFast(const State& st, NrlCD p, NrlCD e) {
    const double parameterSet = (p[99] == 0.0) ? 2.0 : p[99];
    if (parameterSet != 2.0) return -1;
    const double(&plg)[4][9] = st.plg;
    const double cd14 = e[GL_CD14], cd18 = e[GL_CD18], cd32 = e[GL_CD32], cd39 = e[GL_CD39];
    const double t0 = p[21] * st.dfa;
    const double t1 = p[1] * plg[0][2] + p[2] * plg[0][4] + p[22] * plg[0][6] + p[26] * plg[0][1] + p[14] * plg[0][3] + p[59] * plg[0][5];
    const double t2 = (p[18] + p[47] * plg[0][2] + p[29] * plg[0][4]) * cd32;
    const double t3 = (p[15] + p[16] * plg[0][2] + p[30] * plg[0][4]) * cd18;
    const double t4 = (p[9] * plg[0][1] + p[10] * plg[0][3] + p[20] * plg[0][5]) * cd14;
    const double t5 = (p[37] * plg[0][1]) * cd39;
    const double t71 = p[11] * plg[1][2] * cd14;
    const double t72 = p[12] * plg[1][2] * cd14;
    const double t6 = ((p[3] * plg[1][1] + p[4] * plg[1][3] + t71) * st.ctloc + (p[6] * plg[1][1] + p[7] * plg[1][3] + t72) * st.stloc);
    const double t81 = (p[23] * plg[2][3] + p[35] * plg[2][5]) * cd14;
    const double t82 = (p[33] * plg[2][3] + p[36] * plg[2][5]) * cd14;
    const double t7 = ((p[5] * plg[2][2] + p[41] * plg[2][4] + t81) * st.c2tloc + (p[8] * plg[2][2] + p[42] * plg[2][4] + t82) * st.s2tloc);
    const double t13 = p[39] * plg[3][3] * st.s3tloc + p[40] * plg[3][3] * st.c3tloc;
    const double t8 = (p[50] * st.apt[0] + p[96] * plg[0][2] * st.apt[0]);
    const double t10 = (1.0 + plg[0][1] * e[GL_LX] + e[GL_LY]) *
                       ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[74] * plg[1][1] + p[75] * plg[1][3] + p[76] * plg[1][5]) *
                            st.cglong +
                        (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[77] * plg[1][1] + p[78] * plg[1][3] + p[79] * plg[1][5]) *
                            st.sglong);
    double tt = 0;
    tt += t0;
    tt += t1;
    tt += t2;
    tt += t3;
    tt += t4;
    tt += t5;
    tt += t6;
    tt += t7;
    tt += t8;
    tt += t10;
    tt += t13;
    return tt;
}

// Constant phases of a coefficient table's longitude / local-time terms:
// cos and sin of hr p[131], dgtr p[97] and dgtr p[136] (GT_COUNT per table,
// tables in GE order), computed once on the host from the coefficient table.
enum : int { GT_C131 = 0, GT_S131, GT_C97, GT_S97, GT_C136, GT_S136, GT_COUNT };

// Per-object globe7 with the epoch terms E (this table's GE_STRIDE block) and
// the basis of State (plg, local-time and longitude trigonometry). The
// expansion is term by term the published one (Nrlmsise00Model.inl:688-850)
// with cos(x - p) = cos x cos p + sin x sin p for the longitude / local-time
// phases and powers of the Ap decay factor taken by products.
// The Ap function of a globe7 table with p[51] != 0 (the apt[0] its call
// leaves), from the table's epoch terms E.
template <class EP>
__device__ __attribute__((always_inline)) inline double globeApt(NrlCD p, EP E, double glat) {
    if (p[138] == 0.0) return E[GE_APT0];
    double ex = exp(-10800.0 * sqrt(p[51] * p[51]) / (1.0 + p[138] * (45.0 - sqrt(glat * glat))));
    if (ex > 0.99999) ex = 0.99999;
    const auto g = E + GE_G0;
    const double e2 = ex * ex, e3 = e2 * ex, e4 = e2 * e2, e8 = e4 * e4, e12 = e8 * e4, e16 = e8 * e8, e19 = e16 * e3;
    const double sx = 1.0 + (1.0 - e19) / (1.0 - ex) * sqrt(ex);
    return (g[0] + (g[1] * ex + g[2] * e2 + g[3] * e3 + (g[4] * e4 + g[5] * e12) * (1.0 - e8) / (1.0 - ex))) / sx;
}

template <class EP>
__device__ __attribute__((always_inline)) inline double globe7Fast(const State& st, NrlCD p, EP E, double dfa, double glat,
                                                                   double ut71, double ut79c, double ut79s, double ut58,
                                                                   NrlCD tr, double& aptOut) {
    const double(&plg)[4][9] = st.plg;
    const double cd14 = E[GE_CD14], cd18 = E[GE_CD18], f2 = E[GE_F2];
    double tinf = p[30];
    tinf += E[GE_T0];
    tinf += (p[1] * plg[0][2] + p[2] * plg[0][4] + p[22] * plg[0][6]) + (p[14] * plg[0][2]) * dfa + p[26] * plg[0][1];
    tinf += p[18] * E[GE_CD32];
    tinf += (p[15] + p[16] * plg[0][2]) * cd18;
    tinf += E[GE_F1] * (p[9] * plg[0][1] + p[10] * plg[0][3]) * cd14;
    tinf += p[37] * plg[0][1] * E[GE_CD39];
    {
        const double t71 = (p[11] * plg[1][2]) * cd14, t72 = (p[12] * plg[1][2]) * cd14;
        tinf += f2 * ((p[3] * plg[1][1] + p[4] * plg[1][3] + p[27] * plg[1][5] + t71) * st.ctloc +
                      (p[6] * plg[1][1] + p[7] * plg[1][3] + p[28] * plg[1][5] + t72) * st.stloc);
    }
    {
        const double t81 = (p[23] * plg[2][3] + p[35] * plg[2][5]) * cd14, t82 = (p[33] * plg[2][3] + p[36] * plg[2][5]) * cd14;
        tinf += f2 * ((p[5] * plg[2][2] + p[41] * plg[2][4] + t81) * st.c2tloc + (p[8] * plg[2][2] + p[42] * plg[2][4] + t82) * st.s2tloc);
    }
    double apt0 = 0.0;
    if (p[51] != 0) {
        apt0 = globeApt(p, E, glat);
        const double c131 = st.ctloc * tr[GT_C131] + st.stloc * tr[GT_S131];
        tinf += apt0 * (p[50] + p[96] * plg[0][2] + p[54] * plg[0][4] +
                        (p[125] * plg[0][1] + p[126] * plg[0][3] + p[127] * plg[0][5]) * cd14 +
                        (p[128] * plg[1][1] + p[129] * plg[1][3] + p[130] * plg[1][5]) * c131);
    }
    const double cg = st.cglong, sg = st.sglong;
    tinf += (1.0 + p[80] * dfa) *
            ((p[64] * plg[1][2] + p[65] * plg[1][4] + p[66] * plg[1][6] + p[103] * plg[1][1] + p[104] * plg[1][3] + p[105] * plg[1][5] +
              (p[109] * plg[1][1] + p[110] * plg[1][3] + p[111] * plg[1][5]) * cd14) * cg +
             (p[90] * plg[1][2] + p[91] * plg[1][4] + p[92] * plg[1][6] + p[106] * plg[1][1] + p[107] * plg[1][3] + p[108] * plg[1][5] +
              (p[112] * plg[1][1] + p[113] * plg[1][3] + p[114] * plg[1][5]) * cd14) * sg);
    {
        const double c2g = cg * cg - sg * sg, s2g = 2.0 * sg * cg;
        double t11 = (1.0 + p[95] * plg[0][1]) * (1.0 + p[81] * dfa) * (1.0 + p[119] * plg[0][1] * cd14) *
                     ((p[68] * plg[0][1] + p[69] * plg[0][3] + p[70] * plg[0][5]) * ut71);
        t11 += (p[76] * plg[2][3] + p[77] * plg[2][5] + p[78] * plg[2][7]) * (ut79c * c2g - ut79s * s2g) *
               (1.0 + p[137] * dfa);
        tinf += t11;
    }
    if (p[51] != 0) {
        const double c97 = cg * tr[GT_C97] + sg * tr[GT_S97];
        const double c136 = cg * tr[GT_C136] + sg * tr[GT_S136];
        tinf += apt0 * (1. + p[132] * plg[0][1]) * ((p[52] * plg[1][2] + p[98] * plg[1][4] + p[67] * plg[1][6]) * c97) +
                apt0 * (p[133] * plg[1][1] + p[134] * plg[1][3] + p[135] * plg[1][5]) * cd14 * c136 +
                apt0 * (p[55] * plg[0][1] + p[56] * plg[0][3] + p[57] * plg[0][5]) * ut58;
    }
    tinf += f2 * ((p[39] * plg[3][3] + (p[93] * plg[3][4] + p[46] * plg[3][6]) * cd14) * st.s3tloc +
                  (p[40] * plg[3][3] + (p[94] * plg[3][4] + p[48] * plg[3][6]) * cd14) * st.c3tloc);
    aptOut = apt0;
    return tinf;
}

// The trigonometry of the globe basis: sin and cos of the latitude, of the
// local time's hour angle and of the longitude (fastBasis's arguments).
struct BasisTrig {
    double c, s, sl, cl, cg, sg;
};
__device__ __attribute__((always_inline)) inline BasisTrig basisTrig(const Input& in) {
    const double dgtr = 1.74533E-2, hr = 0.2618;
    BasisTrig t;
    t.c = sin(in.g_lat * dgtr);
    t.s = cos(in.g_lat * dgtr);
    t.sl = sin(hr * in.lst);
    t.cl = cos(hr * in.lst);
    t.cg = cos(dgtr * in.g_long);
    t.sg = sin(dgtr * in.g_long);
    return t;
}

// Globe basis of the fast policies from its trigonometry: Legendre functions
// of the latitude, local-time and longitude harmonics (double and triple
// angles by products).
__device__ __attribute__((always_inline)) inline void fastBasisFrom(State& st, const BasisTrig& t, double dfa) {
    const double c = t.c, s = t.s;
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
    // local time harmonics from one sin/cos pair (double and triple angles)
    const double sl = t.sl, cl = t.cl;
    st.stloc = sl;
    st.ctloc = cl;
    st.s2tloc = 2.0 * sl * cl;
    st.c2tloc = cl * cl - sl * sl;
    st.s3tloc = sl * (3.0 - 4.0 * sl * sl);
    st.c3tloc = cl * (4.0 * cl * cl - 3.0);
    st.cglong = t.cg;
    st.sglong = t.sg;
    st.dfa = dfa;
}

__device__ __attribute__((always_inline)) inline void fastBasis(State& st, const Input& in, double dfa) {
    fastBasisFrom(st, basisTrig(in), dfa);
}

// Fast policy: epoch terms from the per-epoch record.
struct FastGlobe {
    const double* E;  // GE_COUNT values
    const double* T;  // GE_TABLES x GT_COUNT table phases
    __device__ __attribute__((always_inline)) void basis(State& st, const Input& in, const Flags&) const { fastBasis(st, in, E[GE_DFA]); }
    __device__ __attribute__((always_inline)) void basisFrom(State& st, const BasisTrig& t) const { fastBasisFrom(st, t, E[GE_DFA]); }
    // globe7 of table j without side effects; apt0 is the table's Ap function
    __device__ __attribute__((always_inline)) double value(const State& st, int j, const double* p, const Input& in,
                                                          double& apt0) const {
        const NrlCD ec = (NrlCD)(E + GE_STRIDE * j);
        return globe7Fast(st, (NrlCD)p, ec, ((NrlCD)E)[GE_DFA], in.g_lat, ec[GE_UT71], ec[GE_UT79C], ec[GE_UT79S], ec[GE_UT58],
                          (NrlCD)(T + GT_COUNT * j), apt0);
    }
    // the Ap function globe7 of table j (p[51] != 0) leaves in apt[0]
    __device__ __attribute__((always_inline)) double apt(int j, const double* p, const Input& in) const {
        return globeApt((NrlCD)p, (NrlCD)(E + GE_STRIDE * j), in.g_lat);
    }
    __device__ __attribute__((always_inline)) double operator()(State& st, int j, const double* p, const Input& in, const Flags&) const {
        // globe7 also leaves apt[0] for glob7s (lower thermosphere)
        double apt0 = 0.0;
        const double v = value(st, j, p, in, apt0);
        if (p[51] != 0) st.apt[0] = apt0;
        return v;
    }
    __device__ __attribute__((always_inline)) double lower(State& st, int j, const double* p, const Input&, const Flags&) const {
        return glob7sFast(st, (NrlCD)p, (NrlCD)(E + GE_COUNT + GL_STRIDE * j));
    }
};


// Weather-slot rows (uneven epochs): per 3-hour UTC slot of the weather table,
// the globe7 terms that are constant over the slot (day-of-year cosines, flux
// factors, Ap functions) and, per table, cos/sin of the UT phases, so a lane
// forms cos(sr (sec - p)) = cos(sr sec) cos(sr p) + sin(sr sec) sin(sr p).
enum : int {
    GS_C71 = 18, GS_S71, GS_C79, GS_S79, GS_C58, GS_S58, GS_STRIDE = 24,
    SLOT_DFA = GE_TABLES * GS_STRIDE,
    SLOT_DOY = SLOT_DFA + 1,
    SLOT_F107 = SLOT_DFA + 2,    // F10.7 (previous day), then F10.7A, then the 7-entry Ap array
    SLOT_F107A = SLOT_F107 + 1,
    SLOT_AP = SLOT_F107 + 2,
    SLOT_VALID = SLOT_AP + 7,
    SLOT_LOWER = SLOT_VALID + 1,   // lowerEpoch terms (GL_COUNT)
    SLOT_SIZE = kNrlSlotSize
};
static_assert(SLOT_LOWER + GL_COUNT <= SLOT_SIZE, "slot row overflow");

__device__ inline void globeSlot(const NrlTables& T, int doy, double f107, double f107A, const double* ap, double* out) {
    const double sr = 7.2722E-5;
    double E[GE_COUNT];
    globeEpoch(T, doy, 0.0, f107, f107A, ap, E);
    out[SLOT_DFA] = E[GE_DFA];
    lowerEpoch(T, doy, out + SLOT_LOWER);
    for (int j = 0; j < GE_TABLES; ++j) {
        const double* p = globeTable(T, j);
        double* o = out + GS_STRIDE * j;
        for (int k = 0; k < GE_STRIDE; ++k) o[k] = E[GE_STRIDE * j + k];
        o[GE_UT71] = o[GE_UT79C] = o[GE_UT79S] = o[GE_UT58] = 0.0;
        o[GS_C71] = cos(sr * p[71]);
        o[GS_S71] = sin(sr * p[71]);
        o[GS_C79] = cos(sr * p[79]);
        o[GS_S79] = sin(sr * p[79]);
        o[GS_C58] = cos(sr * p[58]);
        o[GS_S58] = sin(sr * p[58]);
    }
}

// Slot policy: the lane's weather-slot row S and cos/sin(sr * sec) of its epoch.
struct SlotGlobe {
    const double* S;
    const double* T;  // GE_TABLES x GT_COUNT table phases
    double cs, ss;
    __device__ __attribute__((always_inline)) void basis(State& st, const Input& in, const Flags&) const {
        fastBasis(st, in, S[SLOT_DFA]);
    }
    __device__ __attribute__((always_inline)) void basisFrom(State& st, const BasisTrig& t) const { fastBasisFrom(st, t, S[SLOT_DFA]); }
    // globe7 of table j without side effects; apt0 is the table's Ap function
    __device__ __attribute__((always_inline)) double value(const State& st, int j, const double* p, const Input& in,
                                                          double& apt0) const {
        const double* e = S + GS_STRIDE * j;
        const double ut71 = cs * e[GS_C71] + ss * e[GS_S71];
        const double ut79c = cs * e[GS_C79] + ss * e[GS_S79];
        const double ut79s = ss * e[GS_C79] - cs * e[GS_S79];
        const double ut58 = cs * e[GS_C58] + ss * e[GS_S58];
        return globe7Fast(st, (NrlCD)p, e, S[SLOT_DFA], in.g_lat, ut71, ut79c, ut79s, ut58, (NrlCD)(T + GT_COUNT * j), apt0);
    }
    __device__ __attribute__((always_inline)) double apt(int j, const double* p, const Input& in) const {
        return globeApt((NrlCD)p, S + GS_STRIDE * j, in.g_lat);
    }
    __device__ __attribute__((always_inline)) double operator()(State& st, int j, const double* p, const Input& in,
                                                               const Flags&) const {
        double apt0 = 0.0;
        const double v = value(st, j, p, in, apt0);
        if (p[51] != 0) st.apt[0] = apt0;
        return v;
    }
    __device__ __attribute__((always_inline)) double lower(State& st, int j, const double* p, const Input&, const Flags&) const {
        return glob7sFast(st, (NrlCD)p, (NrlCD)(S + SLOT_LOWER + GL_STRIDE * j));
    }
};

}  // namespace nrlmsise
}  // namespace phx
