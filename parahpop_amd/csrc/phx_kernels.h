// Host-visible launch wrappers of the HIP kernels (no torch types; plain
// device pointers and POD argument structs).
#pragma once

#include <hip/hip_runtime_api.h>

#include "phx_common.h"

namespace phx {

// Orientation unit of a body with a field of its own (ForceModel::rot*), in a
// .brot2 body table on the device (offsets in doubles): type 1 IAU body model
// (RA, Dec, PM polynomials, then nutation-precession amplitudes; `bary` the
// angle polynomials of its barycenter, -1 none) or type 3 binary frame
// (Chebyshev 3-1-3 Euler angles in the ephemeris body-unit layout).
struct RotUnit {
    int type;
    long long off;
    int degRa, degDec, degPm, npRa, npDec, npPm;     // type 1
    long long bary;
    int baryDeg, baryAngles;
    int inertFrame, nIntervals, pdeg;                // type 3
};

struct EpochArgs {
    int count;
    const double* epochs;   // [count] ET seconds
    double* out;            // [count][REC_SIZE]
    // EOP table, SoA: xp[days], yp[days], ut1-utc[days], tai-utc[days]
    const double* eop;
    int eopDays;
    double eopFirst;
    // tides
    int solid, ocean, zeroTide;
    const double* freq;     // rows of 7: 5 Delaunay multipliers, in-phase, out-of-phase
    int freqCount[3];
    const double* waves;    // rows of 8: 6 Doodson multipliers, first term, end term
    int waveCount;
    const double* terms;    // rows of 6: n, m, Cp, Sp, Cm, Sm
    // NRLMSISE-00 weather (13 x days SoA) and coefficient blocks
    const double* nrlCoeff;
    const double* weather;
    int wDays;
    double wStart;
    // Chebyshev ephemeris (nullptr: analytic series): the body-unit data and, per
    // slot (the configured bodies, then the Sun and the Moon for the tides), a
    // row of kEphRow doubles: step count, then {data offset, intervals, degree,
    // sign} per step of the chain to the Earth
    const double* eph;
    const double* ephChain;
    // IERS2000 Earth orientation payload of a .brot2 unit (nullptr: analytic
    // model) and its nutation / ERP node counts
    const double* iers;
    int iersNut, iersErp;
    // leap-second steps of UT1 ([leapCount] ET, then [leapCount] rotation
    // offsets); gridRows 1: the records are interpolation rows, whose rotation
    // is stored without the steps (R3(-theta(et)) R) so that it is smooth
    const double* leaps;
    int leapCount;
    int gridRows;
    // orientation units of the bodies with fields of their own
    const double* brot;
    int nRot;
    RotUnit rot[kMaxRotFields];
};

constexpr int kEphSteps = 6;
constexpr int kEphRow = 1 + 4 * kEphSteps;
constexpr int kEphSlots = kMaxBodies + 2;

int launchEpochRecords(const EpochArgs& a, const ForceModel& fm, hipStream_t stream);

// Weather-slot rows for every 3-hour slot of the weather table (rows =
// wDays * 8, nrlmsise::SLOT_SIZE doubles each).
int launchNrlSlots(const EpochArgs& a, double* out, int rows, hipStream_t stream);

// Uploads the degree-6 Legendre recursion factors used by the tide gradient.
int uploadTideFactors();

// Uploads the constant longitude / local-time phases of the NRLMSISE-00
// globe7 tables, computed from the host copy of the coefficient table.
int uploadGlobeTrig(const double* nrlCoeffHost);

// RKF78 step kernel: `waves` waves per 64-object tile (1, 2 or 4 with drag;
// also 8 or 16 without).
int launchStep(const StepArgs& a, const ForceModel& fm, int waves, hipStream_t stream);

// Split stages (shared-epoch batches): k-stages in HBM, one drag-free launch
// (`shWaves` = 4, 8 or 16 waves per tile) and one drag launch per stage, and a
// combine launch per step.
struct SplitArgs {
    double* ks;      // [13][6][nPad] k-stages
    double* ys;      // [7][nPad] stage position, velocity, epoch
    double* parts;   // [6][nPad] point-mass + SH part and corrections part of the stage acceleration
    double* spec;    // [8][nPad] species densities (drag spread over several workgroups per tile)
    double* lane;    // [5][nPad] NRLMSISE-00 lane inputs of the stage: altitude, latitude, longitude, local time, state
    int stage;       // 0 .. 12
    int step;        // absolute step index (epoch record of (step, stage))
    int speciesPerGroup;  // 0: one four-wave drag workgroup per tile; 1, 2, 4: species per workgroup
    // SH over several workgroups per tile (few active tiles): `groups`
    // workgroups of the SH kernel's waves share the tile's order blocks
    // (groupBlockStart over groups * waves); each writes its partial SH sums to
    // shPart ([groups][3][nPad]) and the first also the other families to
    // famPart ([6][nPad]); the last to arrive (tileCount, reset by it) forms k_s.
    int groups;
    double* shPart;
    double* famPart;
    unsigned* tileCount;
    int groupBlockStart[kMaxGroupWaves + 1];
    // shared-epoch records: the step index read on the device (nullptr: `step`);
    // the combine kernel advances it, so a graph of several steps replays with
    // the same arguments
    int* stepDev;
    // [6][nPad] the globe basis trigonometry of the stage's lane inputs
    // (nrlmsise::BasisTrig), left by the SH kernel for the drag tile kernel
    // (nullptr: the drag tile kernel forms it)
    double* trig;
};
int launchSplitStage(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, int shWaves, hipStream_t stream);
int launchSplitCombine(const StepArgs& a, const SplitArgs& sp, hipStream_t stream);

// Diagnostics used by the unit tests: one force evaluation per object at the
// given states and epoch record (acc out [3][n]); `which` selects a force family
// (-1 all, 0 central, 1 SH, 2 third bodies, 3 SRP, 4 drag, 5 Earth corrections).
int launchForceProbe(int n, const double* pos, const double* vel, const double* ams, const double* bcd,
                     const double* record, int which, const ForceModel& fm, const double* shTerms,
                     const int* shOrderStart, const double* shDiag, const double* nrlCoeff, double* acc,
                     hipStream_t stream);

// NRLMSISE-00 gtd7d on the device for the published test-driver inputs
// (in: rows of 16 = doy, sec, alt, lat, lon, lst, f107a, f107, ap, ap[7];
// out: rows of 11 = d[9], t[2]), switches as in the published driver.
int launchNrlmsiseProbe(int n, const double* in, int useApArray, const double* nrlCoeff, double* out, hipStream_t stream);

}  // namespace phx
