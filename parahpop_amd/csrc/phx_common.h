// Layouts shared by the host C-ABI layer and the HIP kernels.
#pragma once

#include <cstdint>

namespace phx {

constexpr int kStages = 13;
constexpr int kTile = 64;          // objects per workgroup tile = one wave64
constexpr int kMaxBodies = 11;     // Earth, Sun, Moon, 8 planetary system barycenters
constexpr int kNrlSlotSize = 312;  // doubles per NRLMSISE-00 weather-slot row
constexpr int kMaxWaves = 16;      // waves per tile that share one object's work
constexpr int kMaxGroups = 4;      // SH workgroups per tile when few tiles are active
constexpr int kMaxGroupWaves = kMaxGroups * kMaxWaves;
constexpr int kMaxExpSegments = 32; // bands of a piecewise exponential atmosphere
constexpr int kMaxRotFields = 3;    // bodies other than the Earth with a gravity field

// Per-epoch record: everything the force model needs that depends only on the
// stage epoch, computed once per distinct epoch and read by every object.
enum EpochRecord : int {
    REC_R = 0,           // 9: ICRF -> ITRF rotation, row-major
    REC_OMEGA = 9,       // 3: Earth angular velocity in ICRF (rad/s)
    REC_INDIRECT = 12,   // 3: sum over third bodies of -gm b/|b|^3 (km/s^2)
    REC_SPARE = 15,      // 3
    REC_BODY = 18,       // 3 per configured body (config order), geocentric km
    REC_TIDE_C = REC_BODY + 3 * kMaxBodies,   // 28: C[n][m], n <= 6, at n(n+1)/2+m
    REC_TIDE_S = REC_TIDE_C + 28,             // 28
    REC_NRL = REC_TIDE_S + 28,                // NRLMSISE-00 epoch inputs
    REC_NRL_DOY = REC_NRL + 0,
    REC_NRL_SEC = REC_NRL + 1,
    REC_NRL_GAST = REC_NRL + 2,   // hours
    REC_NRL_F107 = REC_NRL + 3,
    REC_NRL_F107A = REC_NRL + 4,
    REC_NRL_AP = REC_NRL + 5,     // 7 entries (daily Ap, then the 3-hour history)
    REC_NRL_VALID = REC_NRL + 12,
    REC_NRL_GLOBE = REC_NRL + 13, // epoch-only globe7 terms of the 11 tables (nrlmsise::GE_COUNT = 199),
                                  // then the glob7s terms of the 5 lower-thermosphere tables (nrlmsise::GL_COUNT = 30)
    REC_USED = REC_NRL_GLOBE + 199 + 30,
    REC_BROT = 352,      // 9 per rotating body (ForceModel::rot*): ICRF -> body axes, row-major
    REC_SIZE = REC_BROT + 9 * kMaxRotFields + 5
};
static_assert(REC_USED <= REC_BROT, "epoch record overflow");

// Packed spherical-harmonic term stream. For order m the terms n = m..L are
// consecutive; each carries what the order-block evaluation consumes for that
// degree, already combined with the degree-only weights so the inner loop is
// pure fused multiply-adds on uniform (scalar-loaded) operands.
enum ShTerm : int {
    SHT_C1 = 0,   // (n+1) C_nm
    SHT_S1 = 1,   // (n+1) S_nm
    SHT_C = 2,    // C_nm
    SHT_S = 3,    // S_nm
    SHT_EC = 4,   // e_nm C_nm, e_nm = sqrt((2n+1)(n^2-m^2)/(2n-1))
    SHT_ES = 5,   // e_nm S_nm
    SHT_A = 6,    // a_{n+1,m}  (column recursion factor to reach degree n+1)
    SHT_B = 7,    // b_{n+1,m}
    SHT_STRIDE = 8
};

// Uniform force-model description passed to the kernels by value.
struct ForceModel {
    int nBodies;                  // configured bodies (config order)
    int bodyNaif[kMaxBodies];
    double bodyGm[kMaxBodies];
    int bodyPointGravity[kMaxBodies];
    double gmEarth;
    int earthPointGravity;
    // spherical harmonics (Earth)
    int hasSh;
    int shDegree, shOrder;
    double shRadius, shGm;
    int shNormalized;
    // radiation (Sun) with shadow
    int srp;
    int srpBody;                  // index of the Sun in the body list
    double srpFactor;             // AU^2 * 4.56e-3
    double sunRadius;
    int nOccluders;
    int occBody[4];               // body indices of occulting bodies (Sun excluded)
    double occRadius[4];
    // J2-only Earth gravity (when no shape field is configured)
    int hasJ2;
    double j2Gm, j2Coef, j2Radius, j2Soi;
    // drag (Earth, NRLMSISE-00)
    int drag;
    double hCutoff, densityScale, earthRadius, flattening;
    // drag (Earth, piecewise exponential atmosphere; shares earthRadius and
    // flattening): band s is {rho0 kg/km^3, h0 km, scale height km, cutoff km},
    // sorted ascending by h0
    int expDrag;
    int nExpSeg;
    double expSeg[kMaxExpSegments][4];
    // Earth corrections
    int tides;                    // solid or ocean active
    int relativity;
    double muEarthSI, radiusSI, clightSI;
    // fields of bodies other than the Earth (spherical harmonics or J2), each in
    // its own axes (record rows REC_BROT + 9 k), zero beyond its sphere of influence
    int nRot;
    int rotBody[kMaxRotFields];   // index of the body in the body list
    int rotSh[kMaxRotFields];     // 1: coefficient field, 0: J2
    int rotDegree[kMaxRotFields], rotOrder[kMaxRotFields], rotNormalized[kMaxRotFields];
    double rotGm[kMaxRotFields], rotRadius[kMaxRotFields], rotSoi[kMaxRotFields], rotJ2[kMaxRotFields];
    const double* rotTerms[kMaxRotFields];     // packed SH term stream of the field (FP64)
    const int* rotOrderStart[kMaxRotFields];
    const double* rotDiag[kMaxRotFields];
};

// kernel arguments travel by value (StepArgs + ForceModel + SplitArgs per launch)
static_assert(sizeof(ForceModel) <= 2048, "force model too large for a kernel argument");

// Step-kernel arguments.
struct StepArgs {
    int nTiles;
    int nPad;                    // padded object count (multiple of kTile)
    double* y;                   // [6][nPad] state, km and km/s
    double* cur;                 // [nPad] current epoch ET s
    const double* end;           // [nPad]
    double* nextDt;              // [nPad]
    uint8_t* status;             // [nPad] bit0 terminated, bit1 end of simulation, bit2 padding lane
    const double* ams;           // [nPad] cr*area/mass (km^2/kg)
    const double* bcd;           // [nPad] cd*area/mass
    const double* records;       // epoch records
    const long long* tileRecord; // exact epochs: [nTiles] record index of this tile's (step 0, stage 0)
    const int* tileList;         // [nTiles] pool tile of each workgroup (nullptr: workgroup i runs tile i)
    int stepBase;                // steps completed before this launch
    int steps;                   // steps to run in this launch
    double step;                 // nominal step (s)
    const double* shTerms;       // packed term stream (FP64)
    const float* shTerms32;      // packed term stream (FP32) for mixed precision
    const int* shOrderStart;     // [L+2] first term of order m
    const double* shDiag;        // [L+1] sectoral factors c_m, then [L+1] products c_1 ... c_m
    int blockStart[kMaxWaves + 1];  // order blocks per wave: [blockStart[w], blockStart[w+1])
    int fp64Degree;              // mixed precision: degrees above this use FP32 terms (-1: all FP64)
    int fp64Order;               // mixed precision: orders up to this stay FP64 at every degree
    const double* nrlCoeff;      // NRLMSISE-00 coefficient blocks (device)
    long long* counters;         // [4] diagnostics: object-steps completed
    // Uneven epochs: each slot reads its own segment of grid records, rows
    // t_k = slotT0 + k / gridInvH at record index slotRecord + k (k < slotRows),
    // interpolated per object (cubic Lagrange, rows k-1..k+2); NRLMSISE-00
    // time inputs are then computed per object from the weather table.
    int grid;
    double gridInvH;
    const long long* slotRecord;
    const double* slotT0;
    const int* slotRows;
    const double* nrlSlots;      // [wDays * 8][nrlmsise::SLOT_SIZE] weather-slot rows (uneven epochs)
    // Grid rows hold the rotation without the UT1 steps of leap seconds; an
    // object applies R3(theta) of its own stage epoch (leaps: [leapCount] ET of
    // each step, then [leapCount] offsets theta after it)
    const double* leaps;
    int leapCount;
    const double* weather;       // 13 x wDays SoA (CelesTrak components)
    int wDays;
    double wStart;
};

}  // namespace phx
