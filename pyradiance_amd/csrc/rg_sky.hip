// gendaymtx on the GPU.  Host: weather input (gen/loadEPW.c), solar
// position (gen/sun.c:31-79) and the per-step Perez set-up of
// gen/gendaymtx.c (main loop :487-608, ComputeSky :711-805, Perez
// parameters and efficacies :1010-1290).  Device: the patch luminances
// (CalcSkyPatchLumin :1292-1322), the horizontal-illuminance normalisation
// (CalcRelHorzIllum :1209-1218), the sun spread over its nearest patches
// (AddDirect :828-872) and the averaged sky (-A).
#include "rg_sky.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <strings.h>

namespace rg {

namespace {

constexpr double kPI = 3.14159265358979323846;
constexpr double WHTEFFICACY = 179.;              // common/color.h:269
constexpr double DC_SolarConstantE = 1367.0;      // W/m^2
constexpr double DC_SolarConstantL = 127.5;       // klux
constexpr double SUN_ANG_DEG = 0.533;
constexpr double FTINY_S = 1e-6;
inline double deg2rad(double d) { return d * (kPI / 180.); }
inline double rad2deg(double r) { return r * (180. / kPI); }

// Perez, Seals, Michalsky 1993, Table 1: a..e coefficients (4 each) per
// sky-clearness category
const double kPerez[8][20] = {
    {1.3525, -0.2576, -0.2690, -1.4366, -0.7670, 0.0007, 1.2734, -0.1233, 2.8000, 0.6004,
     1.2375, 1.0000, 1.8734, 0.6297, 0.9738, 0.2809, 0.0356, -0.1246, -0.5718, 0.9938},
    {-1.2219, -0.7730, 1.4148, 1.1016, -0.2054, 0.0367, -3.9128, 0.9156, 6.9750, 0.1774,
     6.4477, -0.1239, -1.5798, -0.5081, -1.7812, 0.1080, 0.2624, 0.0672, -0.2190, -0.4285},
    {-1.1000, -0.2515, 0.8952, 0.0156, 0.2782, -0.1812, -4.5000, 1.1766, 24.7219, -13.0812,
     -37.7000, 34.8438, -5.0000, 1.5218, 3.9229, -2.6204, -0.0156, 0.1597, 0.4199, -0.5562},
    {-0.5484, -0.6654, -0.2672, 0.7117, 0.7234, -0.6219, -5.6812, 2.6297, 33.3389, -18.3000,
     -62.2500, 52.0781, -3.5000, 0.0016, 1.1477, 0.1062, 0.4659, -0.3296, -0.0876, -0.0329},
    {-0.6000, -0.3566, -2.5000, 2.3250, 0.2937, 0.0496, -5.6812, 1.8415, 21.0000, -4.7656,
     -21.5906, 7.2492, -3.5000, -0.1554, 1.4062, 0.3988, 0.0032, 0.0766, -0.0656, -0.1294},
    {-1.0156, -0.3670, 1.0078, 1.4051, 0.2875, -0.5328, -3.8500, 3.3750, 14.0000, -0.9999,
     -7.1406, 7.5469, -3.4000, -0.1078, -1.0750, 1.5702, -0.0672, 0.4016, 0.3017, -0.4844},
    {-1.0000, 0.0211, 0.5025, -0.5119, -0.3000, 0.1922, 0.7023, -1.6317, 19.0000, -5.0000,
     1.2438, -1.9094, -4.0000, 0.0250, 0.3844, 0.2656, 1.0468, -0.3788, -2.4517, 1.4656},
    {-1.0500, 0.0289, 0.4260, 0.3590, -0.3250, 0.1156, 0.7781, 0.0025, 31.0625, -14.5000,
     -46.1148, 55.3750, -7.2312, 0.4050, 13.3500, 0.6234, 1.5000, -0.6426, 1.8564, 0.5636}};
// Perez et al. 1990, Table 1 (clearness bins) and Table 4 (luminous efficacy a, b, c, d)
const double kClearLo[8] = {1.000, 1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200};
const double kClearHi[8] = {1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200, 12.01};
const double kDiffEff[8][4] = {{97.24, -0.46, 12.00, -8.91},   {107.22, 1.15, 0.59, -3.95},
                               {104.97, 2.96, -5.53, -8.77},   {102.39, 5.59, -13.95, -13.90},
                               {100.71, 5.94, -22.75, -23.74}, {106.42, 3.83, -36.15, -28.83},
                               {141.88, 1.90, -53.24, -14.03}, {152.23, 0.35, -45.27, -7.98}};
const double kDirEff[8][4] = {{57.20, -4.55, -2.98, 117.12}, {98.99, -3.46, -1.21, 12.38},
                              {109.83, -4.90, -1.71, -8.81}, {110.34, -5.84, -1.99, -4.56},
                              {106.36, -3.97, -1.75, -6.16}, {107.19, -1.25, -1.51, -26.73},
                              {105.75, 0.77, -1.26, -34.44}, {101.18, 1.58, -1.10, -8.29}};
const short kMoDa[13] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334, 365};

// Per-step state of the reference's globals
struct Perez {
  double apwc, sun_zenith, sky_brightness, sky_clearness, diff_irrad, dir_irrad, diff_illum, dir_illum;
  int julian;
  double eccentricity() const {
    double da = (julian - 1.0) * (2.0 * kPI / 365.0);
    return 1.00011 + 0.034221 * cos(da) + 0.00128 * sin(da) + 0.000719 * cos(2.0 * da) + 0.000077 * sin(2.0 * da);
  }
  double airmass() const { return 1.0 / (cos(sun_zenith) + 0.15 * pow(93.885 - rad2deg(sun_zenith), -1.253)); }
  double brightness() const { return diff_irrad * airmass() / (DC_SolarConstantE * eccentricity()); }
  double clearness() const {
    double z3 = sun_zenith * sun_zenith * sun_zenith;
    return ((diff_irrad + dir_irrad) / diff_irrad + 1.041 * z3) / (1.0 + 1.041 * z3);
  }
  int category() const {
    int i;
    for (i = 0; i < 8; ++i)
      if (sky_clearness >= kClearLo[i] && sky_clearness < kClearHi[i]) break;
    return i;
  }
  double diff_ratio(int i) const {
    return kDiffEff[i][0] + kDiffEff[i][1] * apwc + kDiffEff[i][2] * cos(sun_zenith) +
           kDiffEff[i][3] * log(sky_brightness);
  }
  double dir_ratio(int i) const {
    double v = kDirEff[i][0] + kDirEff[i][1] * apwc + kDirEff[i][2] * exp(5.73 * sun_zenith - 5.0) +
               kDirEff[i][3] * sky_brightness;
    return v > 0.0 ? v : 0.0;
  }
  void limit(double cmax) {
    if (sky_clearness > cmax) sky_clearness = cmax;
    else if (sky_clearness < 1.0) sky_clearness = 1.0;
    if (sky_brightness < 0.01) sky_brightness = 0.01;
    else if (sky_brightness > 0.6) sky_brightness = 0.6;
  }
  // CalcSkyParamFromIllum(): photometric input iterated back to irradiance
  int from_illum() {
    double test1 = 0.1, test2 = 0.1;
    int counter = 0;
    diff_irrad = diff_illum * DC_SolarConstantE / (DC_SolarConstantL * 1000.0);
    dir_irrad = dir_illum * DC_SolarConstantE / (DC_SolarConstantL * 1000.0);
    sky_brightness = brightness();
    sky_clearness = clearness();
    limit(12.0);
    while ((fabs(diff_irrad - test1) > 10.0 || fabs(dir_irrad - test2) > 10.0) && counter != 5) {
      test1 = diff_irrad;
      test2 = dir_irrad;
      counter++;
      int index = category();
      diff_irrad = diff_illum / diff_ratio(index);
      dir_irrad = dir_ratio(index);
      if (dir_irrad > 0.1) dir_irrad = dir_illum / dir_irrad;
      sky_brightness = brightness();
      sky_clearness = clearness();
      limit(12.0);
    }
    return category();
  }
};

// CalcPerezParam(): Perez 1993 Eqns. 6-8
void perez_params(double* pp, double sz, double eps, double delta, int index) {
  if (eps > 1.065 && eps < 2.8 && delta < 0.2) delta = 0.2;
  double x[5][4];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 4; ++j) x[i][j] = kPerez[index][4 * i + j];
  if (index != 0) {
    for (int i = 0; i < 5; ++i) pp[i] = x[i][0] + x[i][1] * sz + delta * (x[i][2] + x[i][3] * sz);
  } else {
    pp[0] = x[0][0] + x[0][1] * sz + delta * (x[0][2] + x[0][3] * sz);
    pp[1] = x[1][0] + x[1][1] * sz + delta * (x[1][2] + x[1][3] * sz);
    pp[4] = x[4][0] + x[4][1] * sz + delta * (x[4][2] + x[4][3] * sz);
    pp[2] = exp(pow(delta * (x[2][0] + x[2][1] * sz), x[2][2])) - x[2][3];
    pp[3] = -exp(delta * (x[3][0] + x[3][1] * sz)) + x[3][2] + delta * x[3][3];
  }
}

double bright3(const double* c) { return 0.2651 * c[0] + 0.6701 * c[1] + 0.0648 * c[2]; }

struct Record {
  int month = 0, day = 0;   // month 0-11
  float hour = 0;
  float a = 0, b = 0;       // the two irradiance / illuminance columns
  bool has_ab = false;
  float dptemp = 0;
  bool has_dp = false;
  int line = 0;
};

[[noreturn]] void wfail(const std::string& m) { throw std::runtime_error(m); }

bool date_ok(int month, int day, double hour) {
  if (month < 0 || month > 11 || hour < 0 || hour > 24 || day < 1) return false;
  return day <= (month == 1 ? 29 : kMoDa[month + 1] - kMoDa[month]);
}

bool starts_nocase(const std::string& s, const char* p) { return strncasecmp(s.c_str(), p, strlen(p)) == 0; }

// EPWopen()/EPWread() of gen/loadEPW.c for WEA and EPW text
void read_weather(const std::string& text, int& wtype, double& lat, double& lon, double& tz,
                  std::vector<Record>& recs) {
  std::istringstream in(text);
  std::string line;
  int lino = 0;
  auto next = [&](const char* what) {
    if (!std::getline(in, line)) wfail(std::string("header read error at line ") + std::to_string(lino + 1));
    ++lino;
    (void)what;
  };
  next("first");
  auto bad = [&]() { wfail("header format error in " + std::string(wtype ? "WEA" : "EPW") + " file at line " +
                           std::to_string(lino)); };
  float tzf = 0, elev = 0;
  int nperhour = 1;
  if (starts_nocase(line, "place ")) {
    wtype = 1;
    next("latitude");
    if (sscanf(line.c_str(), "latitude %lf", &lat) != 1) bad();
    next("longitude");
    if (sscanf(line.c_str(), "longitude %lf", &lon) != 1) bad();
    lon *= -1.;
    next("time_zone");
    if (sscanf(line.c_str(), "time_zone %f", &tzf) != 1) bad();
    tzf *= -1.f / 15.f;
    next("site_elevation");
    if (sscanf(line.c_str(), "site_elevation %f", &elev) != 1) bad();
    if (lat < -90 || lat > 90 || lon < -180 || lon > 180 || tzf < -12 || tzf > 12 || elev < -1000 || elev > 9999.9) bad();
    next("units");
    short u = 0;
    if (sscanf(line.c_str(), "weather_data_file_units %hd", &u) != 1) bad();
    if (u < 1 || u > 3) wfail("weather tape format error in header");
    wtype = u;
  } else if (starts_nocase(line, "LOCATION,")) {
    wtype = 0;
    const char* cp = line.c_str() + 9;
    for (int k = 0; k < 5; ++k) {           // city, state, country, source, WMO
      cp = strchr(cp, ',');
      if (!cp) bad();
      ++cp;
    }
    if (sscanf(cp, "%lf,%lf,%f,%f", &lat, &lon, &tzf, &elev) != 4) bad();
    if (lat < -90 || lat > 90 || lon < -180 || lon > 180 || tzf < -12 || tzf > 12 || elev < -1000 || elev > 9999.9) bad();
    const char* keys[6] = {"DESIGN CONDITIONS,", "TYPICAL/EXTREME PERIODS,", "GROUND TEMPERATURES,",
                           "HOLIDAYS/DAYLIGHT SAVINGS,", "COMMENTS 1,", "COMMENTS 2,"};
    for (const char* k : keys) {
      next(k);
      if (!starts_nocase(line, k)) bad();
    }
    next("DATA PERIODS,");
    if (!starts_nocase(line, "DATA PERIODS,")) bad();
    const char* c2 = strchr(line.c_str() + 13, ',');
    if (!c2) bad();
    nperhour = atoi(c2 + 1);
    if (nperhour <= 0) bad();
  } else {
    wfail("header format error in EPW file at line 1");
  }
  tz = tzf;
  while (std::getline(in, line)) {
    ++lino;
    if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
    Record r;
    r.line = lino;
    if (wtype) {
      short mo, da;
      float hr;
      int used = 0;
      if (sscanf(line.c_str(), "%hd %hd %f%n", &mo, &da, &hr, &used) != 3) wfail("Unable to scan date in WEA input");
      r.month = mo - 1;
      r.day = da;
      r.hour = hr;
      if (sscanf(line.c_str() + used, "%f %f", &r.a, &r.b) != 2)
        wfail("bad format at line " + std::to_string(lino) + " in WEA input");
      r.has_ab = true;
    } else {
      short yr, mo, da;
      int hr, mi, used = 0;
      if (sscanf(line.c_str(), "%hd,%hd,%hd,%d,%d,%n", &yr, &mo, &da, &hr, &mi, &used) != 5)
        wfail("Unable to scan date in EPW input");
      r.month = mo - 1;
      r.day = da;
      r.hour = (float)(hr - 1);
      r.hour += nperhour == 1 ? .5f : mi * (1.f / 60.f);
      // fields after the date: uncertainty flags, then numbers (loadEPW.c:453-530)
      std::vector<std::string> f;
      std::string rest = line.substr((size_t)used);
      size_t p = 0;
      while (true) {
        size_t q = rest.find(',', p);
        f.push_back(rest.substr(p, q == std::string::npos ? std::string::npos : q - p));
        if (q == std::string::npos) break;
        p = q + 1;
      }
      auto num = [&](size_t k, float& v) {
        if (k >= f.size() || f[k].empty()) return false;
        v = (float)atof(f[k].c_str());
        return true;
      };
      // a field at its missing-data value counts as unset (EPWrecInit, loadEPW.c:31-35)
      r.has_dp = num(2, r.dptemp) && r.dptemp != 99.9f;   // 0 uncert, 1 dry bulb, 2 dew point
      r.a = r.b = 9999.f;
      num(9, r.a);                              // direct normal irradiance
      num(10, r.b);                             // diffuse horizontal irradiance
      r.has_ab = true;
    }
    if (!date_ok(r.month, r.day, r.hour)) wfail("Illegal date/time in " + std::string(wtype ? "WEA" : "EPW") + " input");
    recs.push_back(r);
  }
}

}  // namespace

SkyPlan sky_plan(const std::string& weather, const SkyOpts& o) {
  SkyPlan P;
  P.opts = o;
  if (o.mfactor < 1 || o.mfactor > 64) throw std::runtime_error("bad Reinhart subdivision (-m)");
  int wtype = 0;
  double lat = 0, lon = 0, tz = 0;
  std::vector<Record> recs;
  read_weather(weather, wtype, lat, lon, tz, recs);
  P.latitude = lat;
  P.longitude = lon;
  // site geometry as gen/sun.c holds it: radians, longitude and meridian west
  const double s_lat = deg2rad(lat), s_lon = deg2rad(-lon), s_mer = deg2rad(-15. * tz);
  const int input = wtype == 3 ? 2 : 1;
  const bool dir_is_horiz = wtype == 2;
  // rh_init(): patch 0 ground, then Reinhart rows, last the zenith (float, as there)
  {
    static const int tnaz[7] = {30, 30, 24, 24, 18, 12, 6};
    const int m = o.mfactor;
    const double alpha = (kPI / 2.) / (7 * m + .5);
    int np = 0;
    for (int r = 0; r < 7; ++r) np += tnaz[r];
    np = np * m * m + 2;
    P.npatch = np;
    P.palt.assign(np, 0.f);
    P.pazi.assign(np, 0.f);
    P.dom.assign(np, 0.f);
    P.palt[0] = (float)(-kPI / 2.);
    P.dom[0] = (float)(2. * kPI);
    P.palt[np - 1] = (float)(kPI / 2.);
    P.dom[np - 1] = (float)(2. * kPI * (1. - cos(alpha * .5)));
    int p = 1;
    for (int i = 0; i < 7 * m; ++i) {
      const float ralt = (float)(alpha * (i + .5));
      const int ninrow = tnaz[i / m] * m;
      const float dom = (float)(2. * kPI * (sin(alpha * (i + 1)) - sin(alpha * i)) / (double)ninrow);
      for (int j = 0; j < ninrow; ++j) {
        P.palt[p] = ralt;
        P.pazi[p] = (float)(2. * kPI * j / (double)ninrow);
        P.dom[p++] = dom;
      }
    }
  }
  float grefl[3], skyc[3], sunc[3];
  for (int c = 0; c < 3; ++c) {
    grefl[c] = (float)o.ground_color[c];
    skyc[c] = (float)o.sky_color[c];
    sunc[c] = (float)o.sun_color[c];
  }
  const bool sky_on = bright3(o.sky_color) > 1e-4, sun_on = bright3(o.sun_color) > 1e-4;
  double dew_point = 11.0;
  float dpt0 = -100, dpt1 = 0;
  int leap_day = 0;
  for (const Record& r : recs) {
    const int mo = r.month + 1, da = r.day;
    const double hr = r.hour;
    if (r.has_dp && wtype == 0) {            // 3-hour dew point average
      if (dpt0 < -99) dpt0 = dpt1 = r.dptemp;
      dew_point = (1. / 3.) * (dpt0 + dpt1 + r.dptemp);
      dpt0 = dpt1;
      dpt1 = r.dptemp;
    } else {
      dpt0 = -100;
    }
    Perez S;
    if (mo == 2 && da == 29) {
      S.julian = 60;
      leap_day = 1;
    } else {
      S.julian = kMoDa[mo - 1] + da + leap_day;
    }
    const double sda = 0.4093 * sin((2. * kPI / 365.) * (S.julian - 81));
    const double sta = 0.170 * sin((4. * kPI / 373.) * (S.julian - 80)) - 0.129 * sin((2. * kPI / 355.) * (S.julian - 8)) +
                       (12. / kPI) * (s_mer - s_lon);
    const double st = hr + sta;
    const double altitude = asin(sin(s_lat) * sin(sda) - cos(s_lat) * cos(sda) * cos(st * (kPI / 12.)));
    const bool sun_in_sky = altitude > -deg2rad(SUN_ANG_DEG / 2.);
    if (o.sun_hours_only && !sun_in_sky) continue;
    const double azimuth = -atan2(cos(sda) * sin(st * (kPI / 12.)),
                                  -cos(s_lat) * sin(sda) - sin(s_lat) * cos(sda) * cos(st * (kPI / 12.))) +
                           kPI - deg2rad(o.rotation);
    double dir, dif;
    if (wtype <= 1 && (!r.has_ab || r.a == 9999.f || r.b == 9999.f))
      wfail("missing required irradiances at line " + std::to_string(r.line));
    if (wtype == 2) {                        // direct horizontal + diffuse: global, then minus diffuse
      float glob = r.a + r.b;
      dir = (float)(glob - r.b);
      dif = r.b;
    } else {
      dir = r.a;
      dif = r.b;
    }
    P.ntsteps++;
    const bool first = P.steps.empty();
    const int solar_minute = (24 * 60) * (S.julian - 1) + (int)(hr * 60. + .5);
    SkyStep K;
    memset(&K, 0, sizeof K);
    if (dir + dif <= 1e-4) {                 // night: zero row
      if (!o.average || first) P.steps.push_back(K);
      if (sun_in_sky) P.suns.push_back({solar_minute, 0, altitude, azimuth, 0.0});
      continue;
    }
    if (dir_is_horiz && altitude > FTINY_S) dir /= sin(altitude);
    if (input == 1) {
      S.dir_irrad = dir;
      S.diff_irrad = dif;
    } else {
      S.dir_illum = dir;
      S.diff_illum = dif;
    }
    // ComputeSky() scalars
    S.apwc = exp(0.07 * dew_point - 0.075);
    if (altitude <= 0.0) S.sun_zenith = deg2rad(90.0);
    else if (altitude >= deg2rad(87.0)) S.sun_zenith = deg2rad(3.0);
    else S.sun_zenith = deg2rad(90.0) - altitude;
    int index;
    if (input == 1) {
      S.sky_brightness = S.brightness();
      S.sky_clearness = S.clearness();
      S.limit(11.9);
      index = S.category();
      S.diff_illum = S.diff_irrad * S.diff_ratio(index);
      S.dir_illum = S.dir_irrad * S.dir_ratio(index);
    } else {
      index = S.from_illum();
    }
    if (o.output == 1) {
      S.diff_illum = S.diff_irrad * WHTEFFICACY;
      S.dir_illum = S.dir_irrad * WHTEFFICACY;
    }
    float g = (float)S.diff_illum;             // ground radiance, float as the matrix
    if (altitude > 0) g = (float)(g + S.dir_illum * sin(altitude));
    g = (float)(g * (1. / kPI / WHTEFFICACY));
    for (int c = 0; c < 3; ++c) K.ground[c] = g * grefl[c];
    K.sun_zenith = S.sun_zenith;
    K.azimuth = azimuth;
    K.altitude = altitude;
    K.diff_illum = S.diff_illum;
    K.dir_illum = S.dir_illum;
    K.kind = sky_on ? 2 : 1;
    if (sky_on) perez_params(K.perez, S.sun_zenith, S.sky_clearness, S.sky_brightness, index);
    K.add_sun = (S.dir_illum > 1e-4 && sun_on) ? 1 : 0;
    P.steps.push_back(K);
    if (sun_in_sky) P.suns.push_back({solar_minute, 1, altitude, azimuth, S.dir_illum});
  }
  if (!P.ntsteps) throw std::runtime_error("no valid time steps on input");
  P.nsteps = (int)P.steps.size();
  P.ncols = o.average ? 1 : P.nsteps;
  return P;
}

// ---- device -------------------------------------------------------------------

namespace {

constexpr int SKY_BLOCK = 256;
constexpr int MAXSUN = 4;

struct SkyStepOut {
  double scale;        // norm_diff_illum / WHTEFFICACY
  int32_t uniform;     // zero sky: every patch 1
  int32_t nsun;
  int32_t sunp[MAXSUN];
  float sunv[MAXSUN];
};

// Per-patch terms of the sun-sky angle and of the horizontal-illuminance
// weight, from the patch's (float) altitude and azimuth as rh_init() stores
// them (gen/gendaymtx.c:1221-1264).
struct SkyPatchTab {
  double czp, szp;     // cos, sin of the patch zenith angle
  double sinp, domp;   // sin(altitude), solid angle
  double azi;
  double dir[3];       // rh_vector() direction
};

// CalcRelLuminance() with the sun-sky angle of CalcSkyPatchLumin()
// (gen/gendaymtx.c:1292-1322, :1172-1191); czs, szs: cos and sin of the sun
// zenith angle.  The arithmetic is the reference's, term for term, in double.
__device__ __forceinline__ float patch_lum(const SkyStep& k, double czs, double szs, const SkyPatchTab& t) {
  const double aas = fabs(t.azi - k.azimuth);
  const double sspa = acos(czs * t.czp + szs * t.szp * cos(aas));
  const double* pp = k.perez;
  double l = (1.0 + pp[0] * exp(pp[1] / t.czp)) * (1.0 + pp[2] * exp(pp[3] * sspa) + pp[4] * cos(sspa) * cos(sspa));
  float f = (float)l;
  return f < 0 ? 0.f : f;
}

// rh_vector(): patch direction from its (float) altitude and azimuth
__device__ __forceinline__ void patch_dir(double v[3], float alt, float azi) {
  v[1] = cos((double)alt);
  v[0] = v[1] * sin((double)azi);
  v[1] *= cos((double)azi);
  v[2] = sin((double)alt);
}

// Insert (d, p) into the first n entries of a descending list; equal
// products keep the lower patch first.  Fixed-size loops with the bound as a
// condition, so the list stays in registers.
__device__ __forceinline__ void top_insert(double (&d)[MAXSUN], int (&p)[MAXSUN], int n, double dv, int pv) {
  int pos = MAXSUN;
#pragma unroll
  for (int i = MAXSUN - 1; i >= 0; --i)
    if (i < n && (dv > d[i] || (dv == d[i] && pv < p[i] && p[i] >= 0))) pos = i;
#pragma unroll
  for (int j = MAXSUN - 1; j > 0; --j)
    if (j > pos && j < n) { d[j] = d[j - 1]; p[j] = p[j - 1]; }
#pragma unroll
  for (int i = 0; i < MAXSUN; ++i)
    if (i == pos) { d[i] = dv; p[i] = pv; }
}

// x precedes y in the nearest-patch order: larger product, then lower patch;
// empty entries (patch -1) last.
__device__ __forceinline__ bool sun_before(double xd, int xp, double yd, int yp) {
  return xp >= 0 && (yp < 0 || xd > yd || (xd == yd && xp < yp));
}

__global__ void k_sky_patch(const float* __restrict__ palt, const float* __restrict__ pazi,
                            const float* __restrict__ dom, int np, SkyPatchTab* __restrict__ tab) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double zsa = kPI * 0.5 - (double)palt[p];
  SkyPatchTab t;
  t.czp = cos(zsa);
  t.szp = sin(zsa);
  t.sinp = sin((double)palt[p]);
  t.domp = (double)dom[p];
  t.azi = (double)pazi[p];
  patch_dir(t.dir, palt[p], pazi[p]);
  tab[p] = t;
}

// One block per stored step: every patch's relative luminance (written
// step-major to `lum` for the matrix pass), the normalisation sum, and the
// sun's nearest patches.  The sum is per-thread partial sums over a strided
// patch range, then a tree; the nearest patches are per-thread sorted lists
// merged pairwise up a tree, which selects the same patches as the
// reference's one sequential pass (largest dot products, ties to the lower
// patch).
__global__ void __launch_bounds__(SKY_BLOCK) k_sky_step(const SkyStep* __restrict__ steps, int nsteps,
                                                        const SkyPatchTab* __restrict__ tab, int np, int nsuns,
                                                        double fixed_sa, float* __restrict__ lum,
                                                        SkyStepOut* __restrict__ out) {
  const int s = blockIdx.x;
  if (s >= nsteps) return;
  const SkyStep k = steps[s];
  const int tid = threadIdx.x;
  __shared__ double red[SKY_BLOCK];
  __shared__ double cd[SKY_BLOCK * MAXSUN];
  __shared__ int cp[SKY_BLOCK * MAXSUN];
  double acc = 0.0;
  double nd[MAXSUN];
  int npch[MAXSUN];
  for (int i = 0; i < MAXSUN; ++i) { nd[i] = -1.0; npch[i] = -1; }
  const double czs = cos(k.sun_zenith), szs = sin(k.sun_zenith);
  double sv[3];
  sv[1] = cos(k.altitude);
  sv[0] = sv[1] * sin(k.azimuth);
  sv[1] *= cos(k.azimuth);
  sv[2] = sin(k.altitude);
  float* lrow = lum + (size_t)s * np;
  for (int p = 1 + tid; p < np; p += SKY_BLOCK) {
    const SkyPatchTab& t = tab[p];
    if (k.kind == 2) {
      const float f = patch_lum(k, czs, szs, t);
      lrow[p] = f;
      acc += (double)f * t.sinp * t.domp;
    }
    if (k.add_sun) top_insert(nd, npch, nsuns, t.dir[0] * sv[0] + t.dir[1] * sv[1] + t.dir[2] * sv[2], p);
  }
  red[tid] = acc;
  for (int i = 0; i < MAXSUN; ++i) { cd[tid * MAXSUN + i] = nd[i]; cp[tid * MAXSUN + i] = npch[i]; }
  __syncthreads();
  for (int w = SKY_BLOCK / 2; w > 0; w >>= 1) {
    if (tid < w) {
      red[tid] += red[tid + w];
      if (k.add_sun) {
        // merge the sorted lists of threads tid and tid + w into tid's: each
        // entry's merged rank is its own index plus the entries of the other
        // list that precede it (ranks of empty entries may coincide; they
        // are all alike)
        double ad[MAXSUN], bd[MAXSUN];
        int ap[MAXSUN], bp[MAXSUN];
#pragma unroll
        for (int i = 0; i < MAXSUN; ++i) {
          ad[i] = cd[tid * MAXSUN + i];
          ap[i] = cp[tid * MAXSUN + i];
          bd[i] = cd[(tid + w) * MAXSUN + i];
          bp[i] = cp[(tid + w) * MAXSUN + i];
        }
#pragma unroll
        for (int i = 0; i < MAXSUN; ++i) {
          int ra = i, rb = i;
#pragma unroll
          for (int j = 0; j < MAXSUN; ++j) {
            if (j < nsuns && sun_before(bd[j], bp[j], ad[i], ap[i])) ++ra;
            if (j < nsuns && sun_before(ad[j], ap[j], bd[i], bp[i])) ++rb;
          }
          if (i < nsuns && ra < nsuns) { cd[tid * MAXSUN + ra] = ad[i]; cp[tid * MAXSUN + ra] = ap[i]; }
          if (i < nsuns && rb < nsuns) { cd[tid * MAXSUN + rb] = bd[i]; cp[tid * MAXSUN + rb] = bp[i]; }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    SkyStepOut o;
    memset(&o, 0, sizeof o);
    double norm = red[0];
    o.uniform = 0;
    if (k.kind == 2 && norm <= FTINY_S) {
      o.uniform = 1;
      norm = kPI;
    }
    o.scale = k.kind == 2 ? (k.diff_illum / norm) * (1. / WHTEFFICACY) : 0.0;
    o.nsun = 0;
    if (k.add_sun) {
      double wta[MAXSUN], wtot = 0;
#pragma unroll
      for (int i = MAXSUN - 1; i >= 0; --i)
        if (i < nsuns) wtot += wta[i] = 1. / (1.002 - cd[i]);
#pragma unroll
      // the list holds its patches first and empty entries after them, so
      // the i-th placed sun patch is entry i
      for (int i = 0; i < MAXSUN; ++i) {
        if (i >= nsuns || cp[i] < 0) continue;
        float va = (float)(wta[i] * k.dir_illum / (WHTEFFICACY * wtot));
        va = (float)((double)va / (fixed_sa > 0 ? fixed_sa : (double)tab[cp[i]].domp));
        o.sunp[i] = cp[i];
        o.sunv[i] = va;
        o.nsun = i + 1;
      }
    }
    out[s] = o;
  }
}

// The matrix, [patch][step][rgb], from the step-major luminances: a tile of
// FILL_TS steps x FILL_TP patches is read along patches into LDS and written
// along steps, so both sides move whole contiguous runs.
constexpr int FILL_TS = 64, FILL_TP = 16;

__global__ void __launch_bounds__(SKY_BLOCK) k_sky_fill(const SkyStep* __restrict__ steps, const SkyStepOut* __restrict__ so,
                                                        const float* __restrict__ lum, int nsteps, int np, float3 skyc,
                                                        float3 sunc, float* __restrict__ out) {
  __shared__ float lt[FILL_TP][FILL_TS + 1];
  const int s0 = blockIdx.x * FILL_TS, p0 = blockIdx.y * FILL_TP;
  const int tid = threadIdx.x;
  for (int i = tid; i < FILL_TS * FILL_TP; i += SKY_BLOCK) {
    const int sl = i / FILL_TP, pl = i % FILL_TP;
    const int s = s0 + sl, p = p0 + pl;
    float l = 0.f;
    if (s < nsteps && p >= 1 && p < np && steps[s].kind == 2) l = lum[(size_t)s * np + p];
    lt[pl][sl] = l;
  }
  __syncthreads();
  const int sl = tid % FILL_TS;
  const int s = s0 + sl;
  if (s >= nsteps) return;
  const int kind = steps[s].kind;
  const SkyStepOut o = so[s];
  for (int pl = tid / FILL_TS; pl < FILL_TP; pl += SKY_BLOCK / FILL_TS) {
    const int p = p0 + pl;
    if (p >= np) break;
    float v[3] = {0.f, 0.f, 0.f};
    if (p == 0) {
      for (int c = 0; c < 3; ++c) v[c] = steps[s].ground[c];
    } else if (kind == 2) {
      const float l = o.uniform ? 1.f : lt[pl][sl];
      const float sc = (float)((double)l * o.scale);
      v[0] = sc * skyc.x;
      v[1] = sc * skyc.y;
      v[2] = sc * skyc.z;
    }
    if (p > 0 && kind > 0)
      for (int i = 0; i < o.nsun; ++i)
        if (o.sunp[i] == p) {
          v[0] += o.sunv[i] * sunc.x;
          v[1] += o.sunv[i] * sunc.y;
          v[2] += o.sunv[i] * sunc.z;
        }
    float* d = out + ((size_t)p * nsteps + s) * 3;
    d[0] = v[0];
    d[1] = v[1];
    d[2] = v[2];
  }
}

// -A: the rows summed in step order in FP32, then scaled by 1/ntsteps.
__global__ void k_sky_average(const float* __restrict__ full, int np, int nsteps, double inv, int scale,
                              float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * 3) return;
  const int p = i / 3, c = i % 3;
  float acc = full[(size_t)p * nsteps * 3 + c];
  for (int s = 1; s < nsteps; ++s) acc += full[((size_t)p * nsteps + s) * 3 + c];
  out[i] = scale ? (float)(acc * inv) : acc;
}

void ck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// Device buffers of one sky_matrix() call, laid out in one block (256-byte
// aligned parts): step records, step outputs, patch tables, patch altitude /
// azimuth / solid angle, the luminance plane, and with -A the full matrix
// before averaging.
struct SkyScratch {
  size_t steps, so, tab, alt, azi, dom, lum, full, bytes;
};

static SkyScratch sky_scratch_layout(const SkyPlan& P) {
  const size_t np = (size_t)P.npatch, ns = (size_t)P.nsteps;
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  SkyScratch L;
  size_t o = 0;
  L.steps = o; o = up(o + sizeof(SkyStep) * ns);
  L.so = o;    o = up(o + sizeof(SkyStepOut) * ns);
  L.tab = o;   o = up(o + sizeof(SkyPatchTab) * np);
  L.alt = o;   o = up(o + sizeof(float) * np);
  L.azi = o;   o = up(o + sizeof(float) * np);
  L.dom = o;   o = up(o + sizeof(float) * np);
  L.lum = o;   o = up(o + sizeof(float) * np * ns);
  L.full = o;  if (P.opts.average) o = up(o + sizeof(float) * 3 * np * ns);
  L.bytes = o;
  return L;
}

// The kernels of one call, with their buffers at `base` (layout above).
static void sky_matrix_run(const SkyPlan& P, float* out, hipStream_t st, char* base, const SkyScratch& L) {
  const int np = P.npatch, ns = P.nsteps;
  SkyStep* dsteps = (SkyStep*)(base + L.steps);
  SkyStepOut* dso = (SkyStepOut*)(base + L.so);
  SkyPatchTab* dtab = (SkyPatchTab*)(base + L.tab);
  float* dalt = (float*)(base + L.alt);
  float* dazi = (float*)(base + L.azi);
  float* ddom = (float*)(base + L.dom);
  float* dlum = (float*)(base + L.lum);
  ck(hipMemcpyAsync(dsteps, P.steps.data(), sizeof(SkyStep) * ns, hipMemcpyHostToDevice, st), "upload");
  ck(hipMemcpyAsync(dalt, P.palt.data(), sizeof(float) * np, hipMemcpyHostToDevice, st), "upload");
  ck(hipMemcpyAsync(dazi, P.pazi.data(), sizeof(float) * np, hipMemcpyHostToDevice, st), "upload");
  ck(hipMemcpyAsync(ddom, P.dom.data(), sizeof(float) * np, hipMemcpyHostToDevice, st), "upload");
  k_sky_patch<<<(np + 255) / 256, 256, 0, st>>>(dalt, dazi, ddom, np, dtab);
  ck(hipGetLastError(), "k_sky_patch");
  int nsuns = P.opts.nsuns < 1 ? 1 : (P.opts.nsuns > MAXSUN ? MAXSUN : P.opts.nsuns);
  k_sky_step<<<ns, SKY_BLOCK, 0, st>>>(dsteps, ns, dtab, np, nsuns, P.opts.fixed_sun_sa, dlum, dso);
  ck(hipGetLastError(), "k_sky_step");
  float* full = P.opts.average ? (float*)(base + L.full) : out;
  const SkyOpts& o = P.opts;
  float3 skyc = make_float3((float)o.sky_color[0], (float)o.sky_color[1], (float)o.sky_color[2]);
  float3 sunc = make_float3((float)o.sun_color[0], (float)o.sun_color[1], (float)o.sun_color[2]);
  dim3 grid((ns + FILL_TS - 1) / FILL_TS, (np + FILL_TP - 1) / FILL_TP);
  k_sky_fill<<<grid, SKY_BLOCK, 0, st>>>(dsteps, dso, dlum, ns, np, skyc, sunc, full);
  ck(hipGetLastError(), "k_sky_fill");
  if (P.opts.average) {
    k_sky_average<<<(np * 3 + 255) / 256, 256, 0, st>>>(full, np, ns, 1. / (double)P.ntsteps, P.ntsteps > 1, out);
    ck(hipGetLastError(), "k_sky_average");
  }
  ck(hipStreamSynchronize(st), "gendaymtx");
}

void sky_matrix(const SkyPlan& P, float* out, hipStream_t st) {
  if (P.nsteps <= 0) return;
  const SkyScratch L = sky_scratch_layout(P);
  char* base = nullptr;
  ck(hipMalloc((void**)&base, L.bytes), "hipMalloc");
  try {
    sky_matrix_run(P, out, st, base, L);
  } catch (...) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(base);
    throw;
  }
  (void)hipFree(base);
}

// ---- the allocation-free form --------------------------------------------------
// Callers that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc) get the same kernels with their buffers in
// this fixed device array.  Each GPU holds its own copy until the process
// ends.  In host code `scratch` only names the symbol: its device address, per
// device, comes from hipGetSymbolAddress at the first call on that device.
static __device__ __align__(256) unsigned char scratch[kSkyScratchBytes];

static std::mutex g_scratch_mu;                       // calls share one array per device
static std::unordered_map<int, char*> g_scratch_ptr;   // device -> address of `scratch`

void sky_matrix_noalloc(const SkyPlan& P, float* out, hipStream_t st) {
  if (P.nsteps <= 0) return;
  const SkyScratch L = sky_scratch_layout(P);
  if (L.bytes > kSkyScratchBytes)
    throw std::runtime_error("gendaymtx needs " + std::to_string(L.bytes) +
                             " bytes of device scratch; the allocation-free entry point has " +
                             std::to_string(kSkyScratchBytes) + " (Reinhart MF:4 over 8760 steps fits, without -A)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  int dev = 0;
  ck(hipGetDevice(&dev), "hipGetDevice");
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    ck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    it = g_scratch_ptr.emplace(dev, (char*)p).first;
  }
  sky_matrix_run(P, out, st, it->second, L);
}

}  // namespace rg
