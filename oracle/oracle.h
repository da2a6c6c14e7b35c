/* ORACLE — test infrastructure only (parity checker and timed CPU baseline).
 * C-ABI of the CPU restatement of the reference host propagator
 * (propagators/host/Propagator.h:13-34 + parm/integrate/rk/host/Step.h:109-124).
 * Same argument meaning as phx_propagate in include/phx.h. */
#ifndef PHX_ORACLE_H
#define PHX_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

typedef struct orc_stats {
    double object_steps;      /* completed RKF78 steps summed over objects */
    double rhs_evaluations;   /* force-model evaluations (13 per step) */
    double seconds;           /* wall time of the stepping loop */
    int threads;
} orc_stats;

/* data_dir holds the published tables (tide frequencies, FES2004, NRLMSISE-00). */
int orc_create(const char* config_json_path, const char* data_dir, orc_ctx** out);
int orc_propagate(orc_ctx* ctx, size_t n, const int64_t* id, const double* epoch_mjd2000_tdb,
                  const double* state6_soa, const double* mass_kg, const double* drag_area_m2,
                  const double* srp_area_m2, const double* cd, const double* cr, double duration_s,
                  int threads, double* out_epoch, double* out_state6_soa, orc_stats* stats);
/* One right-hand-side evaluation (acceleration, km/s^2) for unit tests of the
 * force model: r, v in km, km/s; epoch ET seconds; area km^2; cr effective. */
int orc_acceleration(orc_ctx* ctx, const double* r, const double* v, double et, double mass,
                     double area_km2, double cr_eff, double cd, double* acc_out);
/* Selected force terms at one state (km/s^2): which = 0 central, 1 SH, 2 third
 * bodies, 3 SRP (with shadow), 4 drag, 5 Earth corrections. */
int orc_force_term(orc_ctx* ctx, int which, const double* r, const double* v, double et, double mass,
                   double area_km2, double cr_eff, double cd, double* acc_out);
/* Raw NRLMSISE-00 gtd7d call (cgs units, switch 0 off as in the published
 * test driver; use_ap_array selects switch 9 = -1 with the 7-entry ap array). */
int orc_nrlmsise(const char* data_dir, int doy, double sec, double alt, double glat, double glon, double lst,
                 double f107a, double f107, double ap, const double* apa, int use_ap_array, double* d9, double* t2);
/* Geocentric body position (km) from the analytic ephemeris and the analytic
 * Earth orientation (row-major ICRF->ITRF matrix, omega rad/s) at ET seconds,
 * with zero polar motion and UT1 = UTC = TT - 69.184 s. */
int orc_body_position(int naif, double et, double* pos3);
int orc_earth_orientation(double et, double* R9, double* omega3);
/* Geocentric body position (km) from the context's ephemeris: its Chebyshev
 * tables when the configuration names .brie files, else the analytic series. */
int orc_ctx_body_position(const orc_ctx* ctx, int naif, double et, double* pos3);
/* ICRF -> ITRF rotation (row-major) and Earth rate from the context's
 * orientation: its IERS2000 unit when the configuration names a .brot2 file,
 * else the analytic model. */
int orc_ctx_earth_orientation(const orc_ctx* ctx, double et, double* R9, double* omega3);
// ICRF -> axes of the configured field of body `naif` (not the Earth); returns
// the rotation target the body resolved to, -1 when the body has no such field.
int orc_body_axes(const orc_ctx* ctx, int naif, double et, double* R9);
/* Geodetic altitude (km) of r above the ellipsoid (equatorial radius Req km,
 * flattening f) whose polar axis is the unit vector pole, r and pole in one frame. */
double orc_geodetic_altitude(const double* r3, const double* pole3, double Req, double f);
/* Density (kg/km^3) of the configured Earth exponential atmosphere at geodetic
 * altitude h km; NaN when the configuration has none. */
double orc_exp_density(const orc_ctx* ctx, double h);
/* Test hooks: single functions of the restatement (pins).
 * orc_time_at: the EOP time model at ET (mjdUtc, UT1-UTC s, TT-UTC s, xp, yp
 * arcsec, valid). orc_nrl_inputs: NRLMSISE-00 inputs at ET for an Earth-fixed
 * position (mjdUtc, doy, seconds of day, GAST h, local time h, altitude km,
 * geodetic latitude and longitude deg, F10.7, F10.7A, daily Ap, ap[7], valid).
 * orc_solid_tide: solid-tide coefficients (C, S [7][7]) for Moon and Sun
 * positions (m, ICRF), rotation E (row-major ICRF->ITRF) and time data
 * (mjdUtc, UT1-UTC, TT-UTC, xp, yp); frequency_terms 0 leaves out the
 * frequency-dependent corrections. orc_ocean_waves: the FES2004 wave sum for
 * Delaunay arguments (l, l', F, D, Omega) and GMST. orc_tide_arguments: those
 * arguments and the Doodson arguments for time data. orc_third_body: the
 * non-barycentric third-body acceleration. orc_shadow: the dual-cone factor
 * (Sun and occulter relative to the object, km). orc_sh_body_fixed: the SH
 * acceleration at a body-fixed position. orc_ocean_tide: the whole ocean-tide
 * coefficient set (waves and the ocean pole tide) for time data.
 * orc_tide_acceleration: the acceleration (m/s^2, ICRF) of the degree
 * 2..max_degree potential with coefficients C, S at an ICRF position (m). */
int orc_time_at(const orc_ctx* ctx, double et, double* out6);
int orc_nrl_inputs(const orc_ctx* ctx, double et, const double* fixed3, double* out19);
int orc_solid_tide(const orc_ctx* ctx, const double* moon3, const double* sun3, const double* E9, const double* time5,
                   int frequency_terms, int zero_tide, double* C49, double* S49);
int orc_ocean_waves(const orc_ctx* ctx, const double* args5, double gmst, double* C49, double* S49);
int orc_tide_arguments(const double* time5, double* args5, double* gmst, double* beta6);
int orc_third_body(const double* r3, const double* b3, double gm, double* out3);
double orc_shadow(const double* sun3, const double* body3, double rSun, double rBody);
int orc_sh_body_fixed(const orc_ctx* ctx, const double* p3, double* out3);
int orc_ocean_tide(const orc_ctx* ctx, const double* time5, double* C49, double* S49);
int orc_tide_acceleration(const orc_ctx* ctx, const double* r3, const double* E9, const double* C49, const double* S49,
                          int max_degree, double* out3);
double orc_step_seconds(const orc_ctx* ctx);
double orc_duration_seconds(const orc_ctx* ctx);
/* FLOP counts of the calling thread since orc_flop_reset (liboracle_flops.so,
 * the build with flop_count.h; returns -1 in the ordinary build): out9 =
 * arithmetic FLOPs of the per-object force model and step, of the epoch-only
 * inputs and of the NRLMSISE-00 density, then the transcendental calls of the
 * same three categories, then the entries into their scopes. */
int orc_flop_counts(double* out9);
void orc_flop_reset(void);
const char* orc_last_error(void);
void orc_destroy(orc_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
