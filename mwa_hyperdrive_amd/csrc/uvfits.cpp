// uvfits output of modelled visibilities: the vis-simulate writer path (src/io/write/mod.rs:113-260,
// which hands the visibilities to marlu 0.17.1's UvfitsWriter, a dependency outside the reference
// tree). The file is AIPS random-groups FITS:
//   - primary HDU: PCOUNT 7 group parameters (UU, VV, WW in seconds, BASELINE, DATE as
//     whole-day-relative part and remainder, INTTIM) and per group [channel][pol][re, im,
//     weight] float32, pols XX, YY, XY, YX (CRVAL3 -5, CDELT3 -1), one IF, RA, Dec;
//   - an "AIPS AN" binary table: tile names, un-precessed geodetic tile XYZs and the standard
//     antenna columns.
// The UVWs are formed per timestep as the writer does: the tile XYZs precessed to J2000 at the
// aberrated phase centre's J2000 hour angle (precess_time's hadec_j2000), or, without
// precession, at the phase centre's hour angle at the LMST. Host code only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "skymodel.h"
#include "sm_time.h"

namespace {

constexpr double kVelC = 299792458.0;
constexpr double kPi = 3.14159265358979323846;

void put_be32(std::vector<unsigned char> &b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put_f32(std::vector<unsigned char> &b, float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    put_be32(b, v);
}
void put_f64(std::vector<unsigned char> &b, double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}

// 80-character header cards.
struct Header {
    std::string text;
    void card(const std::string &raw) {
        std::string c = raw.substr(0, 80);
        c.resize(80, ' ');
        text += c;
    }
    void kv(const char *key, const std::string &value, const char *comment = nullptr) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "%-8.8s= %s", key, value.c_str());
        std::string c = buf;
        if (comment != nullptr) {
            c.resize(std::max<size_t>(c.size(), 31), ' ');
            c += " / ";
            c += comment;
        }
        card(c);
    }
    void logical(const char *key, bool v, const char *comment = nullptr) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%20s", v ? "T" : "F");
        kv(key, buf, comment);
    }
    void integer(const char *key, long long v, const char *comment = nullptr) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%20lld", v);
        kv(key, buf, comment);
    }
    void real(const char *key, double v, const char *comment = nullptr) {
        // 17 significant digits: every double reads back exactly.
        char buf[40];
        std::snprintf(buf, sizeof buf, "%20.16E", v);
        kv(key, buf, comment);
    }
    void str(const char *key, const std::string &v, const char *comment = nullptr) {
        std::string q = v;
        q.resize(std::max<size_t>(q.size(), 8), ' ');
        std::string esc;
        for (char ch : q) {
            esc += ch;
            if (ch == '\'') esc += '\'';
        }
        kv(key, "'" + esc + "'", comment);
    }
    void end_block() {
        card("END");
        while (text.size() % 2880 != 0) text += ' ';
    }
};

void pad_block(std::vector<unsigned char> &b) {
    while (b.size() % 2880 != 0) b.push_back(0);
}

// Gregorian calendar date of a Julian date at 0h (JD ending in .5).
void calendar(double jd0, int &y, int &m, int &d) {
    const long z = (long)std::floor(jd0 + 0.5);
    const long a = (long)std::floor((z - 1867216.25) / 36524.25);
    const long b = z + 1 + a - a / 4 + 1524;
    const long c = (long)std::floor((b - 122.1) / 365.25);
    const long dd = (long)std::floor(365.25 * c);
    const long e = (long)std::floor((b - dd) / 30.6001);
    d = (int)(b - dd - (long)std::floor(30.6001 * e));
    m = (int)(e < 14 ? e - 1 : e - 13);
    y = (int)(m > 2 ? c - 4716 : c - 4715);
}

// Geocentric (ECEF) position of a geodetic longitude, latitude and height on WGS84.
void ecef(double lon, double lat, double h, double out[3]) {
    const double a = 6378137.0, f = 1.0 / 298.257223563, e2 = f * (2.0 - f);
    const double n = a / std::sqrt(1.0 - e2 * std::sin(lat) * std::sin(lat));
    out[0] = (n + h) * std::cos(lat) * std::cos(lon);
    out[1] = (n + h) * std::cos(lat) * std::sin(lon);
    out[2] = (n * (1.0 - e2) + h) * std::sin(lat);
}

int tai_utc(double gps) { return sm::tai_minus_utc_at(gps); }
double utc_jd(double gps) { return sm::utc_jd_of_gps(gps); }

}  // namespace

struct SmUvfitsWriter {
    std::FILE *f = nullptr;
    std::string path;
    int ntimes = 0, nbl = 0, nchan = 0, ntiles = 0;
    std::vector<int32_t> pairs;        // [nbl][2] tile indices
    std::vector<double> xyz;           // [ntiles][3] un-precessed geodetic XYZ
    std::vector<std::string> names;
    double lon = 0, lat = 0, height = 0, pc_ra = 0, pc_dec = 0, dut1 = 0;
    double freq0 = 0, dfreq = 0, int_time = 0, jd_zero = 0;
    bool precess = true;
    int written = 0;
    std::vector<unsigned char> buf;
};

extern "C" const char *sm_uvfits_create(const char *path, const SmUvfitsParams *p, SmUvfitsWriter **out) {
    if (path == nullptr || p == nullptr || out == nullptr) return "null argument";
    *out = nullptr;
    if (p->num_timesteps < 1 || p->num_baselines < 1 || p->num_chans < 1 || p->num_tiles < 1)
        return "need at least one timestep, baseline, channel and tile";
    if (p->tile_names == nullptr || p->tile_xyzs == nullptr || p->baseline_tiles == nullptr)
        return "null tile names, positions or baseline tiles";
    if (p->num_tiles > 2047) return "more than 2047 tiles";
    for (int b = 0; b < p->num_baselines; ++b) {
        const int a1 = p->baseline_tiles[2 * b], a2 = p->baseline_tiles[2 * b + 1];
        if (a1 < 0 || a2 < 0 || a1 >= p->num_tiles || a2 >= p->num_tiles || a1 > a2)
            return "baseline tiles must be tile indices with tile1 <= tile2";
    }
    SmUvfitsWriter *w = new SmUvfitsWriter();
    w->path = path;
    w->ntimes = p->num_timesteps;
    w->nbl = p->num_baselines;
    w->nchan = p->num_chans;
    w->ntiles = p->num_tiles;
    w->pairs.assign(p->baseline_tiles, p->baseline_tiles + 2 * (size_t)p->num_baselines);
    w->xyz.assign(p->tile_xyzs, p->tile_xyzs + 3 * (size_t)p->num_tiles);
    for (int t = 0; t < p->num_tiles; ++t) w->names.push_back(p->tile_names[t] != nullptr ? p->tile_names[t] : "");
    w->lon = p->array_longitude;
    w->lat = p->array_latitude;
    w->height = p->array_height;
    w->pc_ra = p->phase_centre_ra;
    w->pc_dec = p->phase_centre_dec;
    w->dut1 = p->dut1;
    w->precess = p->apply_precession != 0;
    w->freq0 = p->first_freq_hz;
    w->dfreq = p->freq_res_hz;
    w->int_time = p->int_time_s;
    // The whole-day reference (PZERO5): the 0h UTC that ends the Julian day (noon to noon) of
    // the first timestamp, so an evening timestamp has a negative first DATE parameter
    // (2456860.5 for 2014-07-21T20:10:28 UTC, src/cli/vis_simulate/tests.rs:145).
    const double jd_first = utc_jd(p->first_gps_seconds);
    w->jd_zero = std::floor(jd_first) + 0.5;
    w->f = std::fopen(path, "wb");
    if (w->f == nullptr) {
        delete w;
        return "cannot open the uvfits file for writing";
    }
    // DATE-OBS: the first timestamp's UTC date and time.
    int y, m, d;
    const double jd0 = std::floor(jd_first - 0.5) + 0.5;
    calendar(jd0, y, m, d);
    const double secs = (jd_first - jd0) * 86400.0;
    const int hh = (int)(secs / 3600.0), mm = (int)((secs - 3600.0 * hh) / 60.0);
    char date[48];
    std::snprintf(date, sizeof date, "%04d-%02d-%02dT%02d:%02d:%06.3f", y, m, d, hh, mm, secs - 3600.0 * hh - 60.0 * mm);
    const double centre = w->freq0 + w->dfreq * (double)(w->nchan / 2);
    Header h;
    h.logical("SIMPLE", true, "Standard FITS format");
    h.integer("BITPIX", -32, "Floating point values");
    h.integer("NAXIS", 6);
    h.integer("NAXIS1", 0, "Random groups, not an image");
    h.integer("NAXIS2", 3, "re, im, weight");
    h.integer("NAXIS3", 4, "polarisations");
    h.integer("NAXIS4", w->nchan, "fine channels");
    h.integer("NAXIS5", 1);
    h.integer("NAXIS6", 1);
    h.logical("EXTEND", true, "Tables follow");
    h.logical("GROUPS", true, "Random group UV data");
    h.integer("PCOUNT", 7, "Number of random parameters");
    h.integer("GCOUNT", (long long)w->ntimes * w->nbl, "Number of groups (rows)");
    h.real("EPOCH", 2000.0);
    h.real("BSCALE", 1.0);
    h.real("BZERO", 0.0);
    h.str("BUNIT", "JY");
    h.str("CTYPE2", "COMPLEX");
    h.real("CRVAL2", 1.0);
    h.real("CDELT2", 1.0);
    h.real("CRPIX2", 1.0);
    h.str("CTYPE3", "STOKES");
    h.real("CRVAL3", -5.0, "XX, YY, XY, YX");
    h.real("CDELT3", -1.0);
    h.real("CRPIX3", 1.0);
    h.str("CTYPE4", "FREQ");
    h.real("CRVAL4", centre);
    h.real("CDELT4", w->dfreq);
    h.real("CRPIX4", (double)(w->nchan / 2 + 1));
    h.str("CTYPE5", "RA");
    h.real("CRVAL5", w->pc_ra * 180.0 / kPi);
    h.real("CDELT5", 1.0);
    h.real("CRPIX5", 1.0);
    h.str("CTYPE6", "DEC");
    h.real("CRVAL6", w->pc_dec * 180.0 / kPi);
    h.real("CDELT6", 1.0);
    h.real("CRPIX6", 1.0);
    const char *ptypes[7] = {"UU", "VV", "WW", "BASELINE", "DATE", "DATE", "INTTIM"};
    for (int i = 0; i < 7; ++i) {
        char k[9];
        std::snprintf(k, sizeof k, "PTYPE%d", i + 1);
        h.str(k, ptypes[i]);
        std::snprintf(k, sizeof k, "PSCAL%d", i + 1);
        h.real(k, 1.0);
        std::snprintf(k, sizeof k, "PZERO%d", i + 1);
        h.real(k, i == 4 ? w->jd_zero : 0.0);
    }
    h.str("DATE-OBS", date);
    h.str("OBJECT", p->obs_name != nullptr ? p->obs_name : "Undefined");
    h.str("TELESCOP", "MWA");
    h.str("INSTRUME", "MWA");
    h.real("OBSRA", w->pc_ra * 180.0 / kPi);
    h.real("OBSDEC", w->pc_dec * 180.0 / kPi);
    h.end_block();
    if (std::fwrite(h.text.data(), 1, h.text.size(), w->f) != h.text.size()) {
        std::fclose(w->f);
        delete w;
        return "cannot write the uvfits header";
    }
    *out = w;
    return nullptr;
}

extern "C" const char *sm_uvfits_write_timestep(SmUvfitsWriter *w, double gps_seconds, const JonesF32 *vis_fb,
                                                const float *weights_fb) {
    if (w == nullptr || vis_fb == nullptr) return "null argument";
    if (w->written >= w->ntimes) return "every timestep of the file is already written";
    // Cross UVWs of this timestep (src/io/write/mod.rs:79-90: the writer precesses the
    // un-precessed tile positions itself).
    std::vector<double> xyz = w->xyz;
    double ha, dec;
    if (w->precess) {
        SmPrecessionInfo info;
        if (const char *e = sm_precess_time(w->lon, w->lat, w->pc_ra, w->pc_dec, gps_seconds, w->dut1, &info)) return e;
        if (const char *e = sm_precess_xyz(&info, reinterpret_cast<const XYZ *>(w->xyz.data()),
                                           reinterpret_cast<XYZ *>(xyz.data()), w->ntiles))
            return e;
        ha = info.hadec_j2000.ha;
        dec = info.hadec_j2000.dec;
    } else {
        ha = sm_get_lmst(w->lon, gps_seconds, w->dut1) - w->pc_ra;
        dec = w->pc_dec;
    }
    const double s_ha = std::sin(ha), c_ha = std::cos(ha), s_dec = std::sin(dec), c_dec = std::cos(dec);
    std::vector<double> tu(w->ntiles), tv(w->ntiles), tw(w->ntiles);
    for (int t = 0; t < w->ntiles; ++t) {
        const double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
        tu[t] = s_ha * x + c_ha * y;
        tv[t] = -s_dec * c_ha * x + s_dec * s_ha * y + c_dec * z;
        tw[t] = c_dec * c_ha * x - c_dec * s_ha * y + s_dec * z;
    }
    const double jd = utc_jd(gps_seconds);
    const float date1 = (float)(jd - w->jd_zero);
    const float date2 = (float)((jd - w->jd_zero) - (double)date1);
    const int stride = w->nchan * 4 * 3;
    std::vector<unsigned char> &b = w->buf;
    b.clear();
    b.reserve((size_t)w->nbl * (7 + stride) * 4);
    for (int bl = 0; bl < w->nbl; ++bl) {
        const int a1 = w->pairs[2 * bl], a2 = w->pairs[2 * bl + 1];
        put_f32(b, (float)((tu[a1] - tu[a2]) / kVelC));
        put_f32(b, (float)((tv[a1] - tv[a2]) / kVelC));
        put_f32(b, (float)((tw[a1] - tw[a2]) / kVelC));
        // AIPS baseline number: 256 a1 + a2 (1-based), or 2048 a1 + a2 + 65536 past 255 tiles.
        const double base = w->ntiles > 255 ? 2048.0 * (a1 + 1) + (a2 + 1) + 65536.0 : 256.0 * (a1 + 1) + (a2 + 1);
        put_f32(b, (float)base);
        put_f32(b, date1);
        put_f32(b, date2);
        put_f32(b, (float)w->int_time);
        for (int ch = 0; ch < w->nchan; ++ch) {
            const float *v = reinterpret_cast<const float *>(&vis_fb[(size_t)ch * w->nbl + bl]);
            const float wt = weights_fb != nullptr ? weights_fb[(size_t)ch * w->nbl + bl] : 1.0f;
            // Jones order XX, XY, YX, YY -> file order XX, YY, XY, YX.
            const int order[4] = {0, 3, 1, 2};
            for (int pol = 0; pol < 4; ++pol) {
                put_f32(b, v[2 * order[pol]]);
                put_f32(b, v[2 * order[pol] + 1]);
                put_f32(b, wt);
            }
        }
    }
    if (std::fwrite(b.data(), 1, b.size(), w->f) != b.size()) return "cannot write uvfits groups";
    w->written += 1;
    return nullptr;
}

extern "C" const char *sm_uvfits_finish(SmUvfitsWriter *w) {
    if (w == nullptr) return "null argument";
    const char *err = nullptr;
    if (w->written != w->ntimes) err = "fewer timesteps written than the file declares";
    // Pad the groups to whole blocks.
    const long long data_bytes = (long long)w->ntimes * w->nbl * (7 + w->nchan * 12) * 4;
    std::vector<unsigned char> pad((size_t)((2880 - data_bytes % 2880) % 2880), 0);
    if (!pad.empty() && std::fwrite(pad.data(), 1, pad.size(), w->f) != pad.size()) err = "cannot pad the groups";
    // AIPS AN table.
    double arr[3];
    ecef(w->lon, w->lat, w->height, arr);
    int y, m, d;
    calendar(w->jd_zero, y, m, d);
    char date[32];
    std::snprintf(date, sizeof date, "%04d-%02d-%02d", y, m, d);
    // GMST at 0h UTC of the reference date, in degrees (UT1 - UTC folded in).
    const double gps0 = (w->jd_zero - 2444244.5) * 86400.0 + (tai_utc((w->jd_zero - 2444244.5) * 86400.0) - 19);
    const double gst0 = sm_get_lmst(0.0, gps0, w->dut1) * 180.0 / kPi;
    const int row_bytes = 8 + 24 + 4 + 4 + 4 + 1 + 4 + 4 + 1 + 4 + 4;
    Header h;
    h.str("XTENSION", "BINTABLE", "Binary table extension");
    h.integer("BITPIX", 8);
    h.integer("NAXIS", 2);
    h.integer("NAXIS1", row_bytes, "Bytes per row");
    h.integer("NAXIS2", w->ntiles, "Rows");
    h.integer("PCOUNT", 0);
    h.integer("GCOUNT", 1);
    h.integer("TFIELDS", 11);
    const char *cols[11][3] = {{"ANNAME", "8A", nullptr},   {"STABXYZ", "3D", "METERS"}, {"NOSTA", "1J", nullptr},
                               {"MNTSTA", "1J", nullptr},  {"STAXOF", "1E", "METERS"},  {"POLTYA", "1A", nullptr},
                               {"POLAA", "1E", "DEGREES"}, {"POLCALA", "1E", nullptr},  {"POLTYB", "1A", nullptr},
                               {"POLAB", "1E", "DEGREES"}, {"POLCALB", "1E", nullptr}};
    for (int i = 0; i < 11; ++i) {
        char k[9];
        std::snprintf(k, sizeof k, "TTYPE%d", i + 1);
        h.str(k, cols[i][0]);
        std::snprintf(k, sizeof k, "TFORM%d", i + 1);
        h.str(k, cols[i][1]);
        if (cols[i][2] != nullptr) {
            std::snprintf(k, sizeof k, "TUNIT%d", i + 1);
            h.str(k, cols[i][2]);
        }
    }
    h.str("EXTNAME", "AIPS AN");
    h.integer("EXTVER", 1);
    h.real("ARRAYX", arr[0]);
    h.real("ARRAYY", arr[1]);
    h.real("ARRAYZ", arr[2]);
    h.real("GSTIA0", gst0);
    h.real("DEGPDY", 360.9856473662862);
    h.real("FREQ", w->freq0 + w->dfreq * (double)(w->nchan / 2));
    h.str("RDATE", date);
    h.real("POLARX", 0.0);
    h.real("POLARY", 0.0);
    h.real("UT1UTC", w->dut1);
    h.real("DATUTC", 0.0);
    h.str("TIMSYS", "UTC");
    h.str("ARRNAM", "MWA");
    h.integer("NUMORB", 0);
    h.integer("NOPCAL", 0);
    h.integer("FREQID", -1);
    h.real("IATUTC", (double)tai_utc(gps0));
    h.str("XYZHAND", "RIGHT");
    h.str("FRAME", "ITRF");
    h.end_block();
    if (std::fwrite(h.text.data(), 1, h.text.size(), w->f) != h.text.size()) err = "cannot write the AN header";
    std::vector<unsigned char> rows;
    for (int t = 0; t < w->ntiles; ++t) {
        std::string n = w->names[t];
        n.resize(8, ' ');
        rows.insert(rows.end(), n.begin(), n.end());
        for (int k = 0; k < 3; ++k) put_f64(rows, w->xyz[3 * t + k]);
        put_be32(rows, (uint32_t)(t + 1));
        put_be32(rows, 0);
        put_f32(rows, 0.0f);
        rows.push_back('X');
        put_f32(rows, 0.0f);
        put_f32(rows, 0.0f);
        rows.push_back('Y');
        put_f32(rows, 90.0f);
        put_f32(rows, 0.0f);
    }
    pad_block(rows);
    if (std::fwrite(rows.data(), 1, rows.size(), w->f) != rows.size()) err = "cannot write the AN rows";
    if (std::fclose(w->f) != 0 && err == nullptr) err = "cannot close the uvfits file";
    delete w;
    return err;
}
